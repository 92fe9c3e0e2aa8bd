// vo_sweeps.hpp -- the two separable sweeps of dexel dilation on gfx950.
//
// Pass 1 (per x-line, along y) replaces VoronoiMorpho2D (src/vor3d/Voronoi2D.cpp)
// and pass 2 (per y-line, along x) replaces SeparatePowerMorpho2D
// (src/vor3d/SeparatePower2D.cpp); together they compute
// VoronoiMorphoVorPower::dilation (src/vor3d/VoronoiVorPower.cpp:24-96).
//
// Work decomposition: one 64-lane wavefront per (line, chunk of C output
// columns).  A chunk sweeps forward from C-rint columns before its first
// output column and backward from rint columns after its last one, so its
// result equals that of a whole-line sweep (nothing further than rint
// columns away can reach it).  The forward results of the chunk wait in LDS
// and are united with the backward ones as those are produced, so each pass
// writes its output to HBM exactly once.
//
// Pass 1 state: z-sorted disjoint pieces (lo, hi, source line).  A new line
// paints over older pieces (newest wins, which is what the half-space sweep
// needs: a newer piece is nearer to every later line).  A piece is culled
// when its interval of the lower envelope of w + dist_z^2 on the current line
// is empty (w = squared line offset); envelope breakpoints between
// consecutive pieces are single crossings, so emptiness is a local test.
// Output triple per piece: (lo, hi, dy) with rho = sqrt(R^2 - dy^2).
//
// Pass 2 state: interiors S (z-disjoint, owner = last line reached; the one
// reaching further wins, as SeparatePower2D.cpp:403-454) and endpoint disks P
// (z-sorted power envelope, culled when their power cell closes for good,
// when a neighbour disk contains them, or when the interiors cover their
// slice for the rest of their life).  Output = union of interior pieces and
// disk slices.
#pragma once

#include "vo_wave.hpp"

namespace vo {

// ---------------------------------------------------------------------------
// Pass 1
// ---------------------------------------------------------------------------

template <class T>
struct Pass1Args {
  int nx, ny;
  int rint;      // floor(R): largest line offset that still reaches
  int chunk;     // output columns per wavefront
  int nchunks;   // ceil(ny / chunk)
  const int64_t* in_off;
  const T* in_z;
  // output triples in slot layout: triple j of column (x, y) at
  // (x * caps + j) * ny + y, so that the j-th triples of an x-line are
  // contiguous (see slot_index)
  T* m_lo;
  T* m_hi;
  uint16_t* m_dy;
  int caps;        // slots per column
  int32_t* m_cnt;  // per column, at count_index: number of triples
  int* status;
  // CSR index of the column at (line x, position y) of this logical grid:
  // x * sx + y * sy (sx = 1, sy = nx sweeps along y, the reference's axis;
  // sx = ny, sy = 1 with nx, ny swapped sweeps along x)
  int64_t sx, sy;
};

// Slot layout of the pass-1 triples: blocks of kSlotBlock consecutive
// triples of one column are contiguous, and the blocks of the columns
// (x, y .. y + 63) follow each other (y-minor).  Pass 1 computes 64
// consecutive y of one x-line per wavefront, so a wavefront fills whole cache
// lines itself; pass 2 reads a block per lane with two 16-byte loads per
// array, coalesced across the wavefront.  `caps` is a multiple of kSlotBlock.
constexpr int kSlotBlock = 8;
__device__ __forceinline__ int64_t slot_index(int ny, int caps, int x, int y, int j) {
  return ((int64_t(x) * (caps / kSlotBlock) + j / kSlotBlock) * ny + y) * kSlotBlock + (j % kSlotBlock);
}
__device__ __forceinline__ int64_t count_index(int ny, int x, int y) { return int64_t(x) * ny + y; }
// Summary of the block holding triple j of column (x, y) (same y-minor order).
__device__ __forceinline__ int64_t summary_index(int ny, int caps, int x, int y, int j) {
  return (int64_t(x) * (caps / kSlotBlock) + j / kSlotBlock) * ny + y;
}

// Batched launches: job j owns the global tiles (or columns)
// [base[j], base[j + 1]); returns the j holding t (base[njobs] = total).
__device__ __forceinline__ int find_job(const int64_t* base, int njobs, int64_t t) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <class T, int CAP, int NCAP, int ARENA, int CMAX>
struct Pass1Smem {
  static constexpr int ECAP = 2 * (CAP + CAP);
  T lo[2][CAP];
  T hi[2][CAP];
  int key[2][CAP];
  T nlo[NCAP];
  T nhi[NCAP];
  int nkey[NCAP];
  T ev[ECAP];
  int ec[ECAP];
  T tlo[ECAP];
  T thi[ECAP];
  int tcode[ECAP];
  double w[CAP];
  int flag[CAP];
  // forward results of the chunk
  T alo[ARENA];
  T ahi[ARENA];
  uint16_t ady[ARENA];
  int apos[CMAX];
  int acnt[CMAX];
  // union staging
  int ka[CAP];
  int kb[CAP];
  T ulo[2 * CAP];
  T uhi[2 * CAP];
  int ukey[2 * CAP];
};

// One sweep step on line k: expire, paint the column's segments, cull.
// Returns the new list size in buffer `cur` (may flip `cur`), or -1 on overflow.
template <class T, class S>
__device__ int pass1_step(S& sm, int& cur, int n, int k, int rint, const T* z, int64_t zoff,
                          int64_t zlen, int ncap, int cap, int* status) {
  const int L = lane_id();
  // expire pieces more than rint lines behind
  {
    const int c = cur;
    n = compact(
        n, [&](int i) { return k - sm.key[c][i] <= rint; },
        [&](int i, int j) {
          sm.lo[c ^ 1][j] = sm.lo[c][i];
          sm.hi[c ^ 1][j] = sm.hi[c][i];
          sm.key[c ^ 1][j] = sm.key[c][i];
        });
    cur ^= 1;
  }
  // paint the column's segments (newest wins), a batch of ncap at a time
  const int64_t npairs = zlen / 2;
  for (int64_t b0 = 0; b0 < npairs; b0 += ncap) {
    const int nb = int((npairs - b0) < ncap ? (npairs - b0) : int64_t(ncap));
    for (int i = L; i < nb; i += kWave) {
      sm.nlo[i] = z[zoff + 2 * (b0 + i)];
      sm.nhi[i] = z[zoff + 2 * (b0 + i) + 1];
      sm.nkey[i] = k;
    }
    wave_sync();
    const int c = cur;
    PieceRef<T> A{sm.lo[c], sm.hi[c], sm.key[c]};
    PieceRef<T> B{sm.nlo, sm.nhi, sm.nkey};
    PieceRef<T> O{sm.lo[c ^ 1], sm.hi[c ^ 1], sm.key[c ^ 1]};
    MergeScratch<T> ms{sm.ev, sm.ec, sm.tlo, sm.thi, sm.tcode, S::ECAP};
    n = merge_max(A, n, B, nb, O, cap, ms);
    if (n < 0) {
      raise_status(status, kStatusP1List);
      return -1;
    }
    cur ^= 1;
  }
  // cull pieces whose envelope interval on this line is empty
  for (int round = 0; round < cap && n > 1; ++round) {
    const int c = cur;
    for (int i = L; i < n; i += kWave) {
      const double d = double(k - sm.key[c][i]);
      sm.w[i] = d * d;
    }
    wave_sync();
    bool any = false;
    for (int base = 0; base < n; base += kWave) {
      const int j = base + L;
      bool kill = false;
      if (j < n) {
        const double lo = sm.lo[c][j], hi = sm.hi[c][j], wj = sm.w[j];
        const double below = j > 0 ? site_crossing(sm.lo[c][j - 1], sm.hi[c][j - 1], sm.w[j - 1], lo, hi, wj)
                                   : -INFINITY;
        const double above = j + 1 < n ? site_crossing(lo, hi, wj, sm.lo[c][j + 1], sm.hi[c][j + 1], sm.w[j + 1])
                                       : INFINITY;
        kill = below >= above;
        sm.flag[j] = kill;
      }
      any |= ballot(kill) != 0;
    }
    wave_sync();
    if (!any) break;
    n = compact(
        n, [&](int i) { return !sm.flag[i]; },
        [&](int i, int j) {
          sm.lo[c ^ 1][j] = sm.lo[c][i];
          sm.hi[c ^ 1][j] = sm.hi[c][i];
          sm.key[c ^ 1][j] = sm.key[c][i];
        });
    cur ^= 1;
  }
  return n;
}

template <class T, int CAP, int NCAP, int ARENA, int CMAX>
__global__ __launch_bounds__(kWave) void pass1_kernel(Pass1Args<T> g) {
  using S = Pass1Smem<T, CAP, NCAP, ARENA, CMAX>;
  __shared__ S sm;
  const int L = lane_id();
  const int x = blockIdx.x % g.nx;
  const int ci = blockIdx.x / g.nx;
  const int y0 = ci * g.chunk;
  const int y1 = min(g.ny, y0 + g.chunk);
  if (y0 >= g.ny) return;
  // forward sweep, k = y
  int cur = 0, n = 0, apos = 0;
  bool ok = true;
  for (int y = max(0, y0 - g.rint); y < y1 && ok; ++y) {
    const int64_t col = int64_t(x) * g.sx + int64_t(y) * g.sy;
    const int64_t zo = g.in_off[col], zl = g.in_off[col + 1] - zo;
    n = pass1_step<T>(sm, cur, n, y, g.rint, g.in_z, zo, zl, NCAP, CAP, g.status);
    if (n < 0) {
      ok = false;
      break;
    }
    if (y >= y0) {
      if (apos + n > ARENA) {
        raise_status(g.status, kStatusArenaOverflow);
        ok = false;
        break;
      }
      for (int i = L; i < n; i += kWave) {
        sm.alo[apos + i] = sm.lo[cur][i];
        sm.ahi[apos + i] = sm.hi[cur][i];
        sm.ady[apos + i] = uint16_t(y - sm.key[cur][i]);
      }
      if (L == 0) {
        sm.apos[y - y0] = apos;
        sm.acnt[y - y0] = n;
      }
      apos += n;
      wave_sync();
    }
  }
  if (!ok) return;
  // backward sweep, k = ny-1-y; unite with the forward result and store
  cur = 0;
  n = 0;
  for (int y = min(g.ny - 1, y1 - 1 + g.rint); y >= y0; --y) {
    const int k = g.ny - 1 - y;
    const int64_t col = int64_t(x) * g.sx + int64_t(y) * g.sy;
    const int64_t zo = g.in_off[col], zl = g.in_off[col + 1] - zo;
    n = pass1_step<T>(sm, cur, n, k, g.rint, g.in_z, zo, zl, NCAP, CAP, g.status);
    if (n < 0) return;
    if (y >= y1) continue;
    const int fa = sm.apos[y - y0], fn = sm.acnt[y - y0];
    for (int i = L; i < fn; i += kWave) sm.ka[i] = -int(sm.ady[fa + i]);
    for (int i = L; i < n; i += kWave) sm.kb[i] = -(k - sm.key[cur][i]);
    wave_sync();
    PieceRef<T> A{sm.alo + fa, sm.ahi + fa, sm.ka};
    PieceRef<T> B{sm.lo[cur], sm.hi[cur], sm.kb};
    PieceRef<T> U{sm.ulo, sm.uhi, sm.ukey};
    MergeScratch<T> ms{sm.ev, sm.ec, sm.tlo, sm.thi, sm.tcode, S::ECAP};
    const int nu = merge_max(A, fn, B, n, U, 2 * CAP, ms);
    if (nu < 0 || nu > g.caps) {
      raise_status(g.status, nu < 0 ? kStatusP1Union : kStatusSlotOverflow);
      return;
    }
    for (int i = L; i < nu; i += kWave) {
      const int64_t o = slot_index(g.ny, g.caps, x, y, i);
      g.m_lo[o] = sm.ulo[i];
      g.m_hi[o] = sm.uhi[i];
      g.m_dy[o] = uint16_t(-sm.ukey[i]);
    }
    if (L == 0) g.m_cnt[count_index(g.ny, x, y)] = nu;
  }
}

// ---------------------------------------------------------------------------
// Pass 2
// ---------------------------------------------------------------------------

template <class T>
struct Pass2Args {
  int nx, ny;
  int rint;
  double R2;
  int chunk;
  int nchunks;
  const T* m_lo;  // pass-1 triples, slot layout (see Pass1Args)
  const T* m_hi;
  const uint16_t* m_dy;
  int caps;
  const int32_t* m_cnt;
  T* o_z;            // output arena (endpoints)
  int64_t arena;     // endpoints reserved per chunk
  int64_t* o_start;  // per column
  int64_t* o_cnt;    // per column, in endpoints
  int* status;
  // list mode: recompute only the gather tiles listed in tiles[1 .. tiles[0]]
  // (columns (x, y0 .. y0 + 63), one block of chunk 1 per column), with
  // arenas from o_base on
  const int* tiles;
  int64_t o_base;
  // CSR index of the column at (line x, position y) of this logical grid:
  // x * sx + y * sy (sx = 1, sy = nx sweeps along y, the reference's axis;
  // sx = ny, sy = 1 with nx, ny swapped sweeps along x)
  int64_t sx, sy;
};

template <class T, int SCAP, int PCAP, int NCAP, int OCAP, int FARENA, int CMAX>
struct Pass2Smem {
  static constexpr int ECAP = 2 * (SCAP + NCAP);
  // interiors
  T slo[2][SCAP];
  T shi[2][SCAP];
  int sD[2][SCAP];
  // disks
  T py[2][PCAP];
  int px[2][PCAP];
  int pdy[2][PCAP];
  int pD[2][PCAP];
  int pflag[PCAP];
  double ps[PCAP];  // slice start / suffix min
  double pe[PCAP];  // slice end / prefix max
  // staged triples of the current column batch
  T nlo[NCAP];
  T nhi[NCAP];
  int nD[NCAP];
  int ndy[NCAP];
  T clo[NCAP];  // non-point interiors of the batch
  T chi[NCAP];
  int cD[NCAP];
  T cy[2 * NCAP];  // candidate disks
  int cdy[2 * NCAP];
  int cDD[2 * NCAP];
  // merge scratch
  T ev[ECAP];
  int ec[ECAP];
  T tlo[ECAP];
  T thi[ECAP];
  int tcode[ECAP];
  // per-step output lists
  double ulo[3][OCAP];
  double uhi[3][OCAP];
  double mlo[2 * OCAP];
  double mhi[2 * OCAP];
  // forward outputs of the chunk
  double flo[FARENA];
  double fhi[FARENA];
  int fpos[CMAX];
  int fcnt[CMAX];
};

// Union of two z-sorted disjoint interval lists (touching ones fuse), merged
// through (mlo, mhi) of capacity cap_m.  Writes at most cap_o intervals to
// (olo, ohi) and returns the count, -1 on overflow.
__device__ int union_lists(const double* alo, const double* ahi, int na, const double* blo,
                           const double* bhi, int nb, double* mlo, double* mhi, int cap_m, double* olo,
                           double* ohi, int cap_o) {
  const int L = lane_id();
  const int n = na + nb;
  if (n > cap_m) return -1;
  for (int i = L; i < na; i += kWave) {
    const int r = rank_in([&](int j) { return blo[j]; }, nb, alo[i], false);
    mlo[i + r] = alo[i];
    mhi[i + r] = ahi[i];
  }
  for (int i = L; i < nb; i += kWave) {
    const int r = rank_in([&](int j) { return alo[j]; }, na, blo[i], true);
    mlo[i + r] = blo[i];
    mhi[i + r] = bhi[i];
  }
  wave_sync();
  int ns = 0, nf = 0;
  double carry = -INFINITY;
  for (int base = 0; base < n; base += kWave) {
    const int i = base + L;
    const double lo = i < n ? mlo[i] : INFINITY;
    const double hi = i < n ? mhi[i] : -INFINITY;
    double run = fmax(wave_scan_max(hi), carry);
    const double up = __shfl_up(run, 1);  // every lane takes part in the shuffle
    const double prev_run = L > 0 ? up : carry;
    const double next_lo = i + 1 < n ? mlo[i + 1] : INFINITY;
    const bool start = i < n && (i == 0 || lo > prev_run);
    const bool fin = i < n && next_lo > run;
    const uint64_t ms = ballot(start), mf = ballot(fin);
    const int is = ns + count_below(ms), jf = nf + count_below(mf);
    if (start && is < cap_o) olo[is] = lo;
    if (fin && jf < cap_o) ohi[jf] = run;
    ns += popc(ms);
    nf += popc(mf);
    carry = __shfl(run, kWave - 1);
  }
  wave_sync();
  return ns > cap_o ? -1 : ns;
}

template <class T, class S>
__device__ int pass2_step(S& sm, int& cs, int& ns, int& cp, int& np, int k, int xcol, int ycol, int mcnt,
                          const Pass2Args<T>& g, int ncap, int scap, int pcap, int ocap,
                          double* outlo, double* outhi) {
  const int L = lane_id();
  // expire interiors and disks whose reach ended
  {
    const int c = cs;
    ns = compact(
        ns, [&](int i) { return sm.sD[c][i] >= k; },
        [&](int i, int j) {
          sm.slo[c ^ 1][j] = sm.slo[c][i];
          sm.shi[c ^ 1][j] = sm.shi[c][i];
          sm.sD[c ^ 1][j] = sm.sD[c][i];
        });
    cs ^= 1;
    const int d = cp;
    np = compact(
        np, [&](int i) { return sm.pD[d][i] >= k; },
        [&](int i, int j) {
          sm.py[d ^ 1][j] = sm.py[d][i];
          sm.px[d ^ 1][j] = sm.px[d][i];
          sm.pdy[d ^ 1][j] = sm.pdy[d][i];
          sm.pD[d ^ 1][j] = sm.pD[d][i];
        });
    cp ^= 1;
  }
  // insert the column's triples, a batch at a time
  if (mcnt < 0 || mcnt > g.caps) {
    raise_status(g.status, kStatusSlotOverflow);
    return -1;
  }
  for (int b0 = 0; b0 < mcnt; b0 += ncap) {
    const int nb = min(ncap, mcnt - b0);
    for (int i = L; i < nb; i += kWave) {
      const int64_t o = slot_index(g.ny, g.caps, xcol, ycol, b0 + i);
      const int dy = g.m_dy[o];
      sm.nlo[i] = g.m_lo[o];
      sm.nhi[i] = g.m_hi[o];
      sm.ndy[i] = dy;
      const int t = reach_of(dy, g.R2);
      sm.nD[i] = t < 0 ? k - 1 : k + t;
    }
    wave_sync();
    // interiors of the batch (points carry no interior)
    const int nc = compact(
        nb, [&](int i) { return sm.nlo[i] < sm.nhi[i] && sm.nD[i] >= k; },
        [&](int i, int j) {
          sm.clo[j] = sm.nlo[i];
          sm.chi[j] = sm.nhi[i];
          sm.cD[j] = sm.nD[i];
        });
    if (nc > 0) {
      const int c = cs;
      PieceRef<T> A{sm.slo[c], sm.shi[c], sm.sD[c]};
      PieceRef<T> B{sm.clo, sm.chi, sm.cD};
      PieceRef<T> O{sm.slo[c ^ 1], sm.shi[c ^ 1], sm.sD[c ^ 1]};
      MergeScratch<T> ms{sm.ev, sm.ec, sm.tlo, sm.thi, sm.tcode, S::ECAP};
      ns = merge_max(A, ns, B, nc, O, scap, ms);
      if (ns < 0) {
        raise_status(g.status, kStatusP2Interiors);
        return -1;
      }
      cs ^= 1;
    }
    // candidate disks: both endpoints of every live triple (one for points)
    int nq = 0;
    for (int base = 0; base < nb; base += kWave) {
      const int i = base + L;
      const bool live = i < nb && sm.nD[i] >= k;
      const bool two = live && sm.nlo[i] < sm.nhi[i];
      const uint64_t m1 = ballot(live), m2 = ballot(two);
      const int pos = nq + count_below(m1) + count_below(m2);
      if (live) {
        sm.cy[pos] = sm.nlo[i];
        sm.cdy[pos] = sm.ndy[i];
        sm.cDD[pos] = sm.nD[i];
        if (two) {
          sm.cy[pos + 1] = sm.nhi[i];
          sm.cdy[pos + 1] = sm.ndy[i];
          sm.cDD[pos + 1] = sm.nD[i];
        }
      }
      nq += popc(m1) + popc(m2);
    }
    wave_sync();
    if (np + nq > pcap) {
      raise_status(g.status, kStatusP2Disks);
      return -1;
    }
    // merge candidates into the disk list by z (older first on ties)
    {
      const int d = cp;
      for (int i = L; i < np; i += kWave) {
        const int r = rank_in([&](int j) { return double(sm.cy[j]); }, nq, double(sm.py[d][i]), false);
        sm.py[d ^ 1][i + r] = sm.py[d][i];
        sm.px[d ^ 1][i + r] = sm.px[d][i];
        sm.pdy[d ^ 1][i + r] = sm.pdy[d][i];
        sm.pD[d ^ 1][i + r] = sm.pD[d][i];
      }
      for (int i = L; i < nq; i += kWave) {
        const int r = rank_in([&](int j) { return double(sm.py[d][j]); }, np, double(sm.cy[i]), true);
        sm.py[d ^ 1][i + r] = sm.cy[i];
        sm.px[d ^ 1][i + r] = k;
        sm.pdy[d ^ 1][i + r] = sm.cdy[i];
        sm.pD[d ^ 1][i + r] = sm.cDD[i];
      }
      np += nq;
      cp ^= 1;
      wave_sync();
    }
  }
  // cull disks
  const double R2 = g.R2;
  for (int round = 0; round <= pcap && np > 0; ++round) {
    const int d = cp, c = cs;
    bool any = false;
    for (int base = 0; base < np; base += kWave) {
      const int q = base + L;
      bool kill = false;
      if (q < np) {
        const double yq = sm.py[d][q];
        const int xq = sm.px[d][q], dyq = sm.pdy[d][q], Dq = sm.pD[d][q];
        const double cq = double(k - xq) * double(k - xq) + double(dyq) * double(dyq);
        const double rq = sqrt(fmax(R2 - double(dyq) * double(dyq), 0.0));
        // covered by interiors for the rest of its reach (first round only)
        if (round == 0 && ns > 0) {
          const double h = sqrt(fmax(R2 - cq, 0.0));
          int j = rank_in([&](int i) { return double(sm.slo[c][i]); }, ns, yq, true) - 1;
          if (j >= 0 && yq <= double(sm.shi[c][j]) && sm.sD[c][j] >= Dq) {
            double lo = sm.slo[c][j], hi = sm.shi[c][j];
            for (int u = j, s = 0; s < 8 && lo > yq - h && u > 0 && sm.shi[c][u - 1] == sm.slo[c][u] &&
                                   sm.sD[c][u - 1] >= Dq;
                 --u, ++s)
              lo = sm.slo[c][u - 1];
            for (int u = j, s = 0; s < 8 && hi < yq + h && u + 1 < ns && sm.slo[c][u + 1] == sm.shi[c][u] &&
                                   sm.sD[c][u + 1] >= Dq;
                 ++u, ++s)
              hi = sm.shi[c][u + 1];
            kill = lo <= yq - h && hi >= yq + h;
          }
        }
        // contained in a neighbouring disk (identical disks: the later one goes)
        for (int side = -1; side <= 1 && !kill; side += 2) {
          const int o = q + side;
          if (o < 0 || o >= np) continue;
          const double yo = sm.py[d][o];
          const int xo = sm.px[d][o], dyo = sm.pdy[d][o];
          const double ro = sqrt(fmax(R2 - double(dyo) * double(dyo), 0.0));
          const double gap = ro - rq;
          const double dx = double(xq - xo), dz = yq - yo;
          if (gap >= 0 && gap * gap >= dx * dx + dz * dz) {
            const bool same = gap == 0 && dx == 0 && dz == 0;
            kill = !same || o < q;
          }
        }
        // power cell closed for good
        if (!kill && q > 0 && q + 1 < np) {
          const double yp = sm.py[d][q - 1], yr = sm.py[d][q + 1];
          const int xp = sm.px[d][q - 1], xr = sm.px[d][q + 1];
          const double cp_ = double(k - xp) * double(k - xp) + double(sm.pdy[d][q - 1]) * double(sm.pdy[d][q - 1]);
          const double cr = double(k - xr) * double(k - xr) + double(sm.pdy[d][q + 1]) * double(sm.pdy[d][q + 1]);
          if (yp < yq && yq < yr) {
            const double bpq = (cq - cp_) / (2.0 * (yq - yp)) + 0.5 * (yp + yq);
            const double bqr = (cr - cq) / (2.0 * (yr - yq)) + 0.5 * (yq + yr);
            const double shrink = double(xq - xr) * (yq - yp) - double(xp - xq) * (yr - yq);
            kill = bpq >= bqr && shrink <= 0;
          }
        }
        if (!kill) {
          // an equal-z neighbour that is lower now and stays lower
          for (int side = -1; side <= 1 && !kill; side += 2) {
            const int o = q + side;
            if (o < 0 || o >= np || sm.py[d][o] != sm.py[d][q]) continue;
            const int xo = sm.px[d][o], dyo = sm.pdy[d][o];
            const double co = double(k - xo) * double(k - xo) + double(dyo) * double(dyo);
            kill = xo >= xq && (co < cq || (co == cq && xo > xq));
          }
        }
        sm.pflag[q] = kill;
      }
      any |= ballot(kill) != 0;
    }
    wave_sync();
    if (!any) break;
    np = compact(
        np, [&](int i) { return !sm.pflag[i]; },
        [&](int i, int j) {
          sm.py[d ^ 1][j] = sm.py[d][i];
          sm.px[d ^ 1][j] = sm.px[d][i];
          sm.pdy[d ^ 1][j] = sm.pdy[d][i];
          sm.pD[d ^ 1][j] = sm.pD[d][i];
        });
    cp ^= 1;
  }
  // slice of the line: union of the interiors and of the disk slices
  const int c = cs, d = cp;
  // interiors: fuse touching pieces
  int nsu = 0;
  {
    int nf = 0;
    for (int base = 0; base < ns; base += kWave) {
      const int i = base + L;
      bool start = false, fin = false;
      if (i < ns) {
        start = i == 0 || sm.shi[c][i - 1] < sm.slo[c][i];
        fin = i + 1 == ns || sm.shi[c][i] < sm.slo[c][i + 1];
      }
      const uint64_t ms = ballot(start), mf = ballot(fin);
      const int is = nsu + count_below(ms), jf = nf + count_below(mf);
      if (start && is < ocap) sm.ulo[0][is] = sm.slo[c][i];
      if (fin && jf < ocap) sm.uhi[0][jf] = sm.shi[c][i];
      nsu += popc(ms);
      nf += popc(mf);
    }
  }
  // disks: components of the centre-sorted slices (prefix max of ends,
  // suffix min of starts)
  for (int i = L; i < np; i += kWave) {
    const double dx = double(k - sm.px[d][i]);
    const double h = sqrt(fmax(R2 - dx * dx - double(sm.pdy[d][i]) * double(sm.pdy[d][i]), 0.0));
    sm.ps[i] = double(sm.py[d][i]) - h;
    sm.pe[i] = double(sm.py[d][i]) + h;
  }
  wave_sync();
  {
    double carry = -INFINITY;
    for (int base = 0; base < np; base += kWave) {
      const int i = base + L;
      double v = fmax(wave_scan_max(i < np ? sm.pe[i] : -INFINITY), carry);
      if (i < np) sm.pe[i] = v;
      carry = __shfl(v, kWave - 1);
    }
    carry = INFINITY;
    const int top = ((np + kWave - 1) / kWave) * kWave;
    for (int base = top - kWave; base >= 0; base -= kWave) {
      const int i = base + L;
      double v = fmin(wave_rscan_min(i < np ? sm.ps[i] : INFINITY), carry);
      if (i < np) sm.ps[i] = v;
      carry = __shfl(v, 0);
    }
  }
  wave_sync();
  int npu = 0;
  {
    int nf = 0;
    for (int base = 0; base < np; base += kWave) {
      const int i = base + L;
      bool start = false, fin = false;
      if (i < np) {
        start = i == 0 || sm.pe[i - 1] < sm.ps[i];
        fin = i + 1 == np || sm.pe[i] < sm.ps[i + 1];
      }
      const uint64_t ms = ballot(start), mf = ballot(fin);
      const int is = npu + count_below(ms), jf = nf + count_below(mf);
      if (start && is < ocap) sm.ulo[1][is] = sm.ps[i];
      if (fin && jf < ocap) sm.uhi[1][jf] = sm.pe[i];
      npu += popc(ms);
      nf += popc(mf);
    }
  }
  wave_sync();
  if (nsu > ocap || npu > ocap) {
    raise_status(g.status, kStatusP2Slice);
    return -1;
  }
  const int nu = union_lists(sm.ulo[0], sm.uhi[0], nsu, sm.ulo[1], sm.uhi[1], npu, sm.mlo, sm.mhi, 2 * ocap,
                             outlo, outhi, ocap);
  if (nu < 0) raise_status(g.status, kStatusP2Slice);
  return nu;
}

template <class T, int SCAP, int PCAP, int NCAP, int OCAP, int FARENA, int CMAX>
__global__ __launch_bounds__(kWave) void pass2_kernel(Pass2Args<T> g) {
  using S = Pass2Smem<T, SCAP, PCAP, NCAP, OCAP, FARENA, CMAX>;
  __shared__ S sm;
  const int L = lane_id();
  int y, x0;
  if (g.tiles) {  // gather tile (x, y0 .. y0 + 63): one block per y, chunk 1
    const int tile = g.tiles[1 + blockIdx.x / kGatherTile];
    x0 = tile % g.nx;
    y = (tile / g.nx) * kGatherTile + blockIdx.x % kGatherTile;
    if (y >= g.ny) return;
  } else {
    y = blockIdx.x % g.ny;
    x0 = (blockIdx.x / g.ny) * g.chunk;
  }
  const int x1 = min(g.nx, x0 + g.chunk);
  if (x0 >= g.nx) return;
  const int64_t obase = g.o_base + int64_t(blockIdx.x) * g.arena;
  if (*g.status != 0) return;  // pass 1 did not complete: its arena is not valid
  // forward sweep, k = x
  int cs = 0, ns = 0, cp = 0, np = 0, fpos = 0;
  for (int x = max(0, x0 - g.rint); x < x1; ++x) {
    const int nu = pass2_step<T>(sm, cs, ns, cp, np, x, x, y, g.m_cnt[count_index(g.ny, x, y)], g, NCAP, SCAP, PCAP, OCAP,
                                 sm.ulo[2], sm.uhi[2]);
    if (nu < 0) return;
    if (x >= x0) {
      if (fpos + nu > FARENA) {
        raise_status(g.status, kStatusArenaOverflow);
        return;
      }
      for (int i = L; i < nu; i += kWave) {
        sm.flo[fpos + i] = sm.ulo[2][i];
        sm.fhi[fpos + i] = sm.uhi[2][i];
      }
      if (L == 0) {
        sm.fpos[x - x0] = fpos;
        sm.fcnt[x - x0] = nu;
      }
      fpos += nu;
      wave_sync();
    }
  }
  // backward sweep, k = nx-1-x; unite and store
  cs = 0;
  ns = 0;
  cp = 0;
  np = 0;
  int64_t opos = 0;
  for (int x = min(g.nx - 1, x1 - 1 + g.rint); x >= x0; --x) {
    const int k = g.nx - 1 - x;
    const int64_t col = int64_t(x) * g.sx + int64_t(y) * g.sy;
    const int nu = pass2_step<T>(sm, cs, ns, cp, np, k, x, y, g.m_cnt[count_index(g.ny, x, y)], g, NCAP, SCAP, PCAP, OCAP,
                                 sm.ulo[2], sm.uhi[2]);
    if (nu < 0) return;
    if (x >= x1) continue;
    const int fa = sm.fpos[x - x0], fn = sm.fcnt[x - x0];
    const int nr = union_lists(sm.flo + fa, sm.fhi + fa, fn, sm.ulo[2], sm.uhi[2], nu, sm.mlo, sm.mhi,
                               2 * OCAP, sm.ulo[0], sm.uhi[0], OCAP);
    if (nr < 0 || opos + 2 * nr > g.arena) {
      raise_status(g.status, nr < 0 ? kStatusP2Slice : kStatusP2Arena);
      return;
    }
    for (int i = L; i < nr; i += kWave) {
      g.o_z[obase + opos + 2 * i] = T(sm.ulo[0][i]);
      g.o_z[obase + opos + 2 * i + 1] = T(sm.uhi[0][i]);
    }
    if (L == 0) {
      g.o_start[col] = obase + opos;
      g.o_cnt[col] = 2 * nr;
    }
    opos += 2 * nr;
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// Column-parallel helpers (one thread per column)
// ---------------------------------------------------------------------------

// The dilation's control words cleared in one launch before pass 1: the
// status, the count of tiles the gather gives back, and the element past the
// output lengths (so that their exclusive scan ends with the total).
__global__ void reset_words_kernel(int* status, int* tiles, int64_t* len_end, int* check_word) {
  *status = 0;
  *tiles = 0;
  *len_end = 0;
  if (check_word) {  // the input check pass 1 does (see init_check_words_kernel)
    check_word[0] = 0;
    check_word[1] = 0x7f7f7f7f;
  }
}
// ... and reported in one launch to the host's pinned words (device-mapped
// memory, read after the stream is synchronised): [0] the output total,
// [1] the status, [2] the tiles given back, [3] the input check's two words
// when pass 1 did the check.
__global__ void report_words_kernel(const int64_t* total, const int* status, const int* tiles, const int* check_word,
                                    int64_t* host) {
  host[0] = *total;
  host[1] = int64_t(uint32_t(*status));
  host[2] = int64_t(uint32_t(*tiles));
  if (check_word) host[3] = int64_t(uint64_t(uint32_t(check_word[0])) | uint64_t(uint32_t(check_word[1])) << 32);
  __threadfence_system();
}
// The input check's words before its kernel: no failure yet, and the
// largest job index (atomicMin keeps the smallest failing one).
__global__ void init_check_words_kernel(int* word) {
  word[0] = 0;
  word[1] = 0x7f7f7f7f;
}

// Gather per-column results (start, count) into CSR at precomputed offsets.
// A column whose start is below `interleaved` was written by the gather form
// of pass 2: its interval i is at start + i * kOutStride (the wavefront's
// i-th intervals side by side); the others are contiguous.
constexpr int kOutStride = 2 * kWave;
// Endpoints past `cap` (the destination's length) are not written.
template <class T>
__global__ void gather_kernel(int64_t ncols, const T* src, const int64_t* start, const int64_t* cnt,
                              const int64_t* off, T* dst, int64_t interleaved, int64_t cap) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int64_t s = start[c], o = off[c];
  const int64_t n = cnt[c];
  if (o < 0 || n < 0 || o + n > cap) return;
  if (s < interleaved) {
    for (int64_t i = 0; i < n; i += 2) {
      dst[o + i] = src[s + (i >> 1) * kOutStride];
      dst[o + i + 1] = src[s + (i >> 1) * kOutStride + 1];
    }
  } else {
    for (int64_t i = 0; i < n; ++i) dst[o + i] = src[s + i];
  }
}

// Device-side check of the CompressedVolume invariants the kernels rely on
// (the host-buffer entry point checks the same on the host): offsets
// non-decreasing with even column lengths, endpoints finite and sorted, and,
// for the operations that complement within a z-range, the first and last
// endpoint of every column inside [lo, hi] (the caller widens or narrows the
// range for the operation).  The reference asserts these with vor_assert
// (src/vor3d/Common.h:228, Common.cpp:8-18).  word[0] collects the bits,
// word[1] the smallest failing job.
template <class T>
struct InputCheck {
  const int64_t* off;
  const T* z;
  int range;       // 1: check [lo, hi]
  double lo, hi;
};

template <class T>
__device__ void check_column(const InputCheck<T>& g, int64_t c, int job, int* word) {
  const int64_t s = g.off[c], e = g.off[c + 1];
  int bad = 0;
  if (e < s || ((e - s) & 1)) {
    bad = kInputMalformed;
  } else if (e > s) {
    T prev = g.z[s];
    int ok = isfinite(double(prev)) ? 1 : 0;
    for (int64_t i = s + 1; i < e; ++i) {
      const T v = g.z[i];
      ok = (ok && isfinite(double(v)) && v >= prev) ? 1 : 0;
      prev = v;
    }
    if (!ok)
      bad = kInputMalformed;
    else if (g.range && (double(g.z[s]) < g.lo || double(prev) > g.hi))
      bad = kInputOutOfRange;
  }
  if (bad) {
    atomicOr(word, bad);
    atomicMin(word + 1, job);
  }
}

// One volume: a wavefront checks 64 consecutive columns, whose endpoints are
// one contiguous range, reading it with coalesced loads 64 endpoints at a
// time, four such loads in flight together (the kernel waits on memory
// latency, not bandwidth); the columns starting inside each chunk are one
// 64-bit mask (an
// OR over the lanes), so that an endpoint is compared with the one before it
// only within its column.  For the ranged operations every endpoint is
// checked against [lo, hi] (with the columns sorted that is the same as
// checking their first and last).
template <class T>
__global__ __launch_bounds__(256) void check_input_kernel(InputCheck<T> g, int64_t ncols, int* word) {
  const int L = lane_id() & (kWave - 1);
  const int64_t c0 = (int64_t(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave) * kWave;
  if (c0 >= ncols) return;
  const int64_t c = c0 + L;
  const bool has = c < ncols;
  const int64_t s = has ? g.off[c] : 0, e = has ? g.off[c + 1] : 0;
  int bad = (has && (e < s || ((e - s) & 1))) ? kInputMalformed : 0;
  // the wavefront's range of endpoints; with malformed offsets only the
  // per-column checks run
  const int last = int(min<int64_t>(ncols - c0, kWave)) - 1;
  const int64_t S = __shfl(s, 0), E = __shfl(e, last);
  if (ballot(bad != 0) == 0 && E > S) {
    T carry = T(0);
    // four chunks of 64 endpoints loaded together (one memory round trip)
    constexpr int U = 4;
    for (int64_t sup = S; sup < E; sup += U * kWave) {
      T vs[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = sup + u * kWave + L;
        vs[u] = i < E ? g.z[i] : T(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t base = sup + u * kWave;
        if (base >= E) break;  // uniform
        const int64_t i = base + L;
        const bool in = i < E;
        const T v = vs[u];
        // bit k: an endpoint base + k starts a (non-empty) column
        const uint64_t mine = (has && s < e && s >= base && s < base + kWave) ? (uint64_t(1) << (s - base)) : 0;
        uint64_t starts = mine;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) starts |= __shfl_xor(starts, d);
        T pv = __shfl_up(v, 1);
        pv = L == 0 ? carry : pv;
        const bool first = ((starts >> L) & 1) != 0 || i == S;
        int b = 0;
        if (in) {
          if (!isfinite(double(v)) || (!first && v < pv)) b = kInputMalformed;
          else if (g.range && (double(v) < g.lo || double(v) > g.hi)) b = kInputOutOfRange;
        }
        bad |= b;
        carry = __shfl(v, kWave - 1);
      }
    }
  }
  if (bad) {
    atomicOr(word, bad);
    atomicMin(word + 1, 0);
  }
}

// Batched: job j owns the columns [col_base[j], col_base[j + 1]).
template <class T>
__global__ void check_input_batch_kernel(const InputCheck<T>* jobs, const int64_t* col_base, int njobs, int* word) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= col_base[njobs]) return;
  const int j = find_job(col_base, njobs, c);
  check_column(jobs[j], c - col_base[j], j, word);
}

// Export of the pass-1 intermediate (the reference's CompressedVolumeWithRadii,
// src/vor3d/CompressedVolumeWithRadii.h:10-38: per column the z-disjoint
// triples (y1, y2, r)) from the slot layout into CSR column order x + nx*y.
// Phase 0 writes each column's triple count, phase 1 the triples at the
// scanned offsets, with r = sqrt(R^2 - dy^2) and the endpoints plus `shift`.
template <class T>
struct TripleExport {
  int nx, ny, caps;
  const T* m_lo;
  const T* m_hi;
  const uint16_t* m_dy;
  const int32_t* m_cnt;
  double R2, shift;
  int64_t* len;
  const int64_t* off;
  double* y1;
  double* y2;
  double* r;
};

template <class T>
__global__ void export_triples_kernel(TripleExport<T> g, int phase) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= int64_t(g.nx) * g.ny) return;
  const int x = int(c % g.nx), y = int(c / g.nx);
  const int n = min(g.m_cnt[count_index(g.ny, x, y)], g.caps);
  if (phase == 0) {
    g.len[c] = n;
    return;
  }
  const int64_t o = g.off[c];
  for (int j = 0; j < n; ++j) {
    const int64_t s = slot_index(g.ny, g.caps, x, y, j);
    const double dy = double(g.m_dy[s]);
    g.y1[o + j] = double(g.m_lo[s]) + g.shift;
    g.y2[o + j] = double(g.m_hi[s]) + g.shift;
    g.r[o + j] = sqrt(fmax(g.R2 - dy * dy, 0.0));
  }
}

// Complement on a grid grown by a ring of full columns
// (VoronoiMorpho::negate, src/vor3d/Voronoi.cpp:18-55, negate_ray
// MorphologyOperators.cpp:230-259).  Phase 0 writes lengths, phase 1 data.
template <class T>
struct ColumnJob {  // one volume of a column-parallel kernel
  int nx, ny;        // input grid
  const int64_t* off;
  const T* z;
  T lo, hi;
  int64_t* out_len;       // per output column
  const int64_t* out_off;  // per output column (absolute into out)
  T* out;
};

template <class T>
__device__ void ring_column(const ColumnJob<T>& g, int64_t C, int phase) {
  const int nx = g.nx, ny = g.ny;
  const int64_t* off = g.off;
  const T* z = g.z;
  const T lo = g.lo, hi = g.hi;
  int64_t* out_len = g.out_len;
  const int64_t* out_off = g.out_off;
  T* out = g.out;
  const int NX = nx + 2;
  const int X = int(C % NX), Y = int(C / NX);
  int64_t s = 0, n = 0;
  if (X >= 1 && X <= nx && Y >= 1 && Y <= ny) {
    const int64_t c = int64_t(X - 1) + int64_t(nx) * (Y - 1);
    s = off[c];
    n = off[c + 1] - s;
  }
  if (n == 0) {
    if (phase == 0) {
      out_len[C] = 2;
    } else {
      out[out_off[C]] = lo;
      out[out_off[C] + 1] = hi;
    }
    return;
  }
  const bool drop_first = z[s] == lo, drop_last = z[s + n - 1] == hi;
  if (phase == 0) {
    out_len[C] = n + (drop_first ? -1 : 1) + (drop_last ? -1 : 1);
    return;
  }
  int64_t w = out_off[C];
  if (!drop_first) out[w++] = lo;
  for (int64_t i = drop_first ? 1 : 0; i < (drop_last ? n - 1 : n); ++i) out[w++] = z[s + i];
  if (!drop_last) out[w++] = hi;
}

template <class T>
__global__ void ring_complement_kernel(ColumnJob<T> g, int phase) {
  const int64_t C = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (C >= int64_t(g.nx + 2) * (g.ny + 2)) return;
  ring_column(g, C, phase);
}

// Crop the ring and complement within [lo, hi] with parity counting
// (negateInv, Voronoi.cpp:57-89; negate_ray_range, MorphologyOperators.cpp:282-315).
// (here nx, ny are the output grid: the input is (nx + 2) x (ny + 2))
template <class T>
__device__ void crop_column(const ColumnJob<T>& g, int64_t c, int phase) {
  const int nx = g.nx;
  const int64_t* off = g.off;
  const T* z = g.z;
  const T lo = g.lo, hi = g.hi;
  int64_t* out_len = g.out_len;
  const int64_t* out_off = g.out_off;
  T* out = g.out;
  const int x = int(c % nx), y = int(c / nx);
  const int64_t C = int64_t(x + 1) + int64_t(nx + 2) * (y + 1);
  const int64_t s = off[C], n = off[C + 1] - s;
  if (n == 0) {
    if (phase == 0) {
      out_len[c] = 2;
    } else {
      out[out_off[c]] = lo;
      out[out_off[c] + 1] = hi;
    }
    return;
  }
  int64_t head = 0;
  while (head < n && z[s + head] <= lo) ++head;
  int64_t tail = 0;
  while (tail < n - head && z[s + n - 1 - tail] >= hi) ++tail;
  const bool add_lo = head % 2 == 0, add_hi = tail % 2 == 0;
  const int64_t mid = n - head - tail;
  if (phase == 0) {
    out_len[c] = mid + (add_lo ? 1 : 0) + (add_hi ? 1 : 0);
    return;
  }
  int64_t w = out_off[c];
  if (add_lo) out[w++] = lo;
  for (int64_t i = head; i < n - tail; ++i) out[w++] = z[s + i];
  if (add_hi) out[w++] = hi;
}

template <class T>
__global__ void crop_complement_kernel(ColumnJob<T> g, int phase) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= int64_t(g.nx) * g.ny) return;
  crop_column(g, c, phase);
}

// Batched forms: job j owns output columns [col_base[j], col_base[j + 1]).
template <class T>
__global__ void ring_complement_batch_kernel(const ColumnJob<T>* jobs, const int64_t* col_base, int njobs,
                                             int phase) {
  const int64_t C = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (C >= col_base[njobs]) return;
  const int j = find_job(col_base, njobs, C);
  ring_column(jobs[j], C - col_base[j], phase);
}

template <class T>
__global__ void crop_complement_batch_kernel(const ColumnJob<T>* jobs, const int64_t* col_base, int njobs,
                                             int phase) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= col_base[njobs]) return;
  const int j = find_job(col_base, njobs, c);
  crop_column(jobs[j], c - col_base[j], phase);
}

// Final offsets of a batch: job j's offsets off_j[0 .. ncols_j] (absolute
// into its endpoint buffer) rebased to start at 0 into rel[out_base[j] ..];
// first[j] = off_j[0] and count[j] = off_j[ncols_j] - off_j[0].
struct FinalJob {
  const int64_t* off;
  int64_t ncols;
  int64_t out_base;
};

// Endpoint count of every job's volume, off[ncols] - off[0].
__global__ void job_counts_kernel(const FinalJob* jobs, int njobs, int64_t* count) {
  const int j = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  const FinalJob g = jobs[j];
  count[j] = g.ncols > 0 ? g.off[g.ncols] - g.off[0] : 0;
}

// n 32-bit words set to v: what the batch program enqueues where a memset
// would go, so that the program can be captured in a HIP graph.
__global__ void fill_kernel(uint32_t* p, int64_t n, uint32_t v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void finalize_offsets_kernel(const FinalJob* jobs, const int64_t* entry_base, int njobs, int64_t* rel,
                                        int64_t* first, int64_t* count) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= entry_base[njobs]) return;
  const int j = find_job(entry_base, njobs, i);
  const FinalJob g = jobs[j];
  const int64_t c = i - entry_base[j];
  const int64_t o0 = g.off[0];
  rel[g.out_base + c] = g.off[c] - o0;
  if (c == 0) {
    first[j] = o0;
    count[j] = g.off[g.ncols] - o0;
  }
}


}  // namespace vo

// ---------------------------------------------------------------------------
// Pass 2, gather form.  The output at column x is the union over source
// columns x' with |x - x'| = d <= R of the pass-1 pieces dilated by
// H = sqrt(R^2 - dy^2 - d^2): exactly what SeparatePower2D's forward and
// backward sweeps assemble (their getLine emits [y1 - dz, y2 + dz] for every
// live interior and point, SeparatePower2D.cpp:295-341).  One wavefront per
// tile of 64 output columns (x, y0 .. y0 + 63), one lane per column.  A block
// of 8 pieces of source column (x + d, y) is read with 16-byte loads, 2 KB
// contiguous per wavefront (the slot layout keeps the blocks of consecutive
// y together), dilated with H from a table (in LDS for R <= 31), and united
// into the lane's sorted output list, kept in global memory with its current
// interval cached in registers.  A tile whose lists overflow makes the
// gather run again with more room, then falls back to the sweep form.
// ---------------------------------------------------------------------------

namespace vo {

template <class T>
struct Pass2GArgs {
  int nx, ny;
  int rint;
  double R2;
  const T* m_lo;
  const T* m_hi;
  const uint16_t* m_dy;
  int caps;
  const int32_t* m_cnt;
  T* o_z;
  int64_t arena;  // endpoints reserved per tile
  int64_t* o_start;
  int64_t* o_cnt;
  int* tile_list;  // [0] = count, then the ids of tiles to recompute with the sweep form
  int* status;
  int ocap;             // intervals per output column list
  const double* htab;   // H by (|d|, dy): htab[|d| * htab_width(rint) + dy], NaN out of reach
  int64_t o_base;       // first endpoint of this volume's tiles in o_z (batched launches)
  double* lists;        // interval lists, 2 * ocap * 64 doubles per tile (lane-minor)
  const T* s_lo;        // block summaries from pass 1 (null: none, no block is skipped)
  const T* s_hi;
  const T* c_hi;        // per column: hi of its last triple (with the summaries)
  // CSR index of the column at (line x, position y) of this logical grid:
  // x * sx + y * sy (sx = 1, sy = nx sweeps along y, the reference's axis;
  // sx = ny, sy = 1 with nx, ny swapped sweeps along x)
  int64_t sx, sy;
};


// Unites [s, e] into the lane's sorted disjoint list (touching fuses),
// starting the search at cursor k; on return k indexes the interval that now
// holds [s, e].  Returns false on overflow.
// The lists are lane-minor: entry k of this lane is at [k * 64] of the
// lane's base pointers olo, ohi.
__device__ __forceinline__ bool lane_unite(double* olo, double* ohi, int ocap, int& n, int& k, double s,
                                           double e) {
  while (k > 0 && ohi[(k - 1) * kWave] >= s) --k;
  while (k < n && ohi[k * kWave] < s) ++k;
  if (k == n || olo[k * kWave] > e) {
    if (n == ocap) return false;
    for (int j = n; j > k; --j) {
      olo[j * kWave] = olo[(j - 1) * kWave];
      ohi[j * kWave] = ohi[(j - 1) * kWave];
    }
    olo[k * kWave] = s;
    ohi[k * kWave] = e;
    ++n;
    return true;
  }
  if (s < olo[k * kWave]) olo[k * kWave] = s;
  double hi = fmax(ohi[k * kWave], e);
  int j = k + 1;
  while (j < n && olo[j * kWave] <= hi) {
    hi = fmax(hi, ohi[j * kWave]);
    ++j;
  }
  const int gone = j - k - 1;
  if (gone > 0) {
    for (int t = k + 1; t + gone < n; ++t) {
      olo[t * kWave] = olo[(t + gone) * kWave];
      ohi[t * kWave] = ohi[(t + gone) * kWave];
    }
    n -= gone;
  }
  ohi[k * kWave] = hi;
  return true;
}

// 8 consecutive values from a 16-byte aligned address, as 16-byte loads.
__device__ __forceinline__ void load_block(const float* p, float (&v)[kSlotBlock]) {
  const float4 a = *reinterpret_cast<const float4*>(p), b = *reinterpret_cast<const float4*>(p + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void load_block(const double* p, double (&v)[kSlotBlock]) {
#pragma unroll
  for (int i = 0; i < kSlotBlock; i += 2) {
    const double2 a = *reinterpret_cast<const double2*>(p + i);
    v[i] = a.x;
    v[i + 1] = a.y;
  }
}

// Source columns whose ranges the gather checks per memory round trip: one
// for R <= 31 (measured fastest there: the prefetched ranges hold registers
// the fifth wave per SIMD needs), four above (97 columns per tile at R = 48).
constexpr int gather_group(bool small_r) { return small_r ? 1 : 4; }

// Largest radius whose H table is staged in LDS ((R + 1)(R + 2) doubles,
// 8.25 KB); above it the lanes read the table from global memory
// (L1-resident).
constexpr int kGatherLdsTableMax = 31;
// Largest radius the gather form takes (its H table is 33.6 MB there); above
// it pass 2 runs as the sweep form, whose LDS lists bound the density it
// takes.  The gather itself keeps its lists in global memory and has no such
// bound (a tile whose lists outgrow their room goes to the sweep form).
constexpr int kGatherMaxR = 2048;

// The gather's H table: row |d| holds H(|d|, dy) = sqrt(R^2 - d^2 - dy^2)
// for dy = 0 .. rint, NaN where that is negative, and one more NaN entry
// (dy = rint + 1) that slots without a triple are pointed at.
__host__ __device__ constexpr int htab_width(int rint) { return rint + 2; }
extern __shared__ double vo_htab_lds[];

// min / max of two doubles as one v_min_f64 / v_max_f64 each.  With one
// operand a quiet NaN they return the other operand; every NaN the gather
// makes is quiet (the table's, and sums with it).  Written as instructions
// because the library calls (fmin / fmax) add canonicalising maxima of loop-
// carried operands on every use.
__device__ __forceinline__ double qmin(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double qmax(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// One tile = (x, y0 .. y0 + 63); consecutive tiles are consecutive x, whose
// source windows overlap, so the launches hand them to one XCD (xcd_tile).
template <class T, bool kLdsTable>
__device__ void gather_tile(const Pass2GArgs<T>& g, int64_t tile) {
  const int L = lane_id();
  const int ocap = g.ocap;
  // the lists are touched only when a run does not meet the cached interval,
  // so they live in global memory (L1/L2-resident) and the kernel needs no
  // LDS: occupancy is set by registers alone
  double* const lbase = g.lists + tile * (2 * int64_t(ocap) * kWave);
  double* olo = lbase + L;
  double* ohi = lbase + ocap * kWave + L;
  const int x = int(tile % g.nx);
  const int y0 = int(tile / g.nx) * kGatherTile;
  if (y0 >= g.ny) return;
  const int rint = g.rint;
  const int y = y0 + L;
  const bool active = y < g.ny;
  const int64_t col = int64_t(x) * g.sx + int64_t(y) * g.sy;
  // every column's output length is valid even where this tile gives up (an
  // overflow, or pass 1 incomplete): the scan and the CSR copy that follow
  // before the host looks at the status then read nothing for it
  if (active) {
    g.o_start[col] = 0;
    g.o_cnt[col] = 0;
  }
  if (*g.status != 0) return;  // pass 1 did not complete
  const double* htab = g.htab;
  if constexpr (kLdsTable) {
    for (int i = L; i < (rint + 1) * htab_width(rint); i += kWave) vo_htab_lds[i] = g.htab[i];
    wave_sync();
    htab = vo_htab_lds;
  }
  const int c_lo = max(0, x - rint), c_hi = min(g.nx - 1, x + rint);
  const int yc = active ? y : g.ny - 1;  // lanes past the grid read a valid row and ignore it
  auto count_of = [&](int c) {
    const int m = min(g.m_cnt[count_index(g.ny, c, yc)], g.caps);
    return active ? m : 0;
  };
  int n = 0;
  // loop-carried lane flags are 0/1 integers: in vector registers they need
  // no exec-mask merges at the joins of the loops
  int ok = 1;
  // Interval k of the list is cached in registers as [clo, chi], with the
  // bounds of its neighbours (below = hi of k - 1, above = lo of k + 1): a run
  // that meets it without reaching a neighbour is united in registers, and
  // only the other runs go through the list in memory.
  int k = 0;
  double clo = INFINITY, chi = -INFINITY, below = -INFINITY, above = INFINITY;
  int dirty = 0;
  // runs parked at the top of the lane's list area (entries ocap - 1 down to
  // ocap - sp); the list itself uses the entries below them
  int sp = 0;
  auto unite_slow = [&](double s, double e) {
    if (dirty) {
      olo[k * kWave] = clo;
      ohi[k * kWave] = chi;
      dirty = 0;
    }
    if (!lane_unite(olo, ohi, ocap - sp, n, k, s, e)) {
      ok = 0;
      return;
    }
    clo = olo[k * kWave];
    chi = ohi[k * kWave];
    below = k > 0 ? ohi[(k - 1) * kWave] : -INFINITY;
    above = k + 1 < n ? olo[(k + 1) * kWave] : INFINITY;
  };
  // unites [s, e] where `want` holds; written with selects so that the
  // wavefront branches only when some lane needs the list
  // A run that needs the list is parked at the top of the lane's list area
  // (two stores), and the parked runs of all lanes go through the list
  // together at the end of the source column, or before a block that could
  // fill a lane's area, so that lanes needing the list at different triples
  // share passes of the wavefront through that path and the list code
  // appears only where runs are flushed.  The union does not depend on the
  // order of its pieces, and a block is skipped only when it lies inside the
  // cached interval, which is part of the union whatever is still parked.
  auto flush_held = [&]() {
    for (;;) {
      const bool h = sp > 0;
      if (ballot(h) == 0) break;
      double s = 0, e = 0;
      if (h) {  // the most recently parked run
        --sp;
        s = olo[(ocap - 1 - sp) * kWave];
        e = ohi[(ocap - 1 - sp) * kWave];
      }
      if (h && ok) unite_slow(s, e);
    }
  };
  auto unite = [&](bool want, double s, double e) {
    if (ballot(want) == 0) return;  // the usual case inside a run
    const bool fast = s <= chi && e >= clo && s > below && e < above;
    const bool f = want && fast;
    clo = f && s < clo ? s : clo;
    chi = f && e > chi ? e : chi;
    dirty = f ? 1 : dirty;
    const bool slow = want && !fast;
    if (ballot(slow) != 0) {
      if (slow) {
        if (n + sp < ocap) {
          olo[(ocap - 1 - sp) * kWave] = s;
          ohi[(ocap - 1 - sp) * kWave] = e;
          ++sp;
        } else {
          ok = 0;
        }
      }
    }
  };
  constexpr int B = kSlotBlock;
  // The lane's block of B contiguous triples of column c from jb on (jb is a
  // multiple of B and below caps, so the whole block is inside the column's
  // slots; values past the column's count are read and discarded): dilated
  // and fused into the open run, which is united when a gap opens.  The
  // wavefront branches only when some lane's run ends.
  struct Block {
    T va[B], vb[B];
    uint4 q;  // the B 16-bit line offsets
  };
  auto load_item = [&](int c, int jb, Block& bk) {
    const int64_t o = slot_index(g.ny, g.caps, c, yc, jb);
    load_block(g.m_lo + o, bk.va);
    load_block(g.m_hi + o, bk.vb);
    bk.q = *reinterpret_cast<const uint4*>(g.m_dy + o);
  };
  // A lane's open run [rs, re] starts as (NaN, NaN).  Slots without a triple
  // (past the column's count, or every slot of a block the lane does not
  // need) read the table's NaN column, as do triples out of reach: their s
  // and e are NaN, no comparison with them holds, and qmin / qmax return the
  // other operand, so they change nothing and need no test of their own.
  const double qnan = __builtin_nan("");
  const int hsent = rint + 1;
  auto compute_block = [&](const Block& bk, int jb, int m, const double* hrow, double& rs, double& re) {
    const T* va = bk.va;
    const T* vb = bk.vb;
    const int nv = ok ? min(max(m - jb, 0), B) : 0;  // the block's slots holding triples
    // the 16-bit offsets of slots u >= nv set to 0xffff, which clamps to hsent
    const uint64_t q0 = (uint64_t(bk.q.y) << 32 | bk.q.x) | (nv >= 4 ? 0 : ~uint64_t(0) << (16 * nv));
    const uint64_t q1 = (uint64_t(bk.q.w) << 32 | bk.q.z) |
                        (nv >= 8 ? 0 : nv <= 4 ? ~uint64_t(0) : ~uint64_t(0) << (16 * (nv - 4)));
    auto hread = [&](int u) { return hrow[min(int(((u < 4 ? q0 : q1) >> (16 * (u & 3))) & 0xffffu), hsent)]; };
    // with FP32 storage and R > 31 the block's table reads go first, in
    // flight together (measured: no gain at R <= 31, where the registers
    // go to a fifth wave, and FP64 storage has none left)
    constexpr bool kAhead = sizeof(T) == 4 && !kLdsTable;
    double hv[B];
    if constexpr (kAhead) {
#pragma unroll
      for (int u = 0; u < B; ++u) hv[u] = hread(u);
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const double h = kAhead ? hv[u] : hread(u);
      const double s = double(va[u]) - h, e = double(vb[u]) + h;
      // triples are z-sorted, so e >= rs whenever s <= re: the run stays
      // connected until a triple starts above it
      const bool gap = s > re;
      if (ballot(gap) != 0) {
        unite(gap, rs, re);
        rs = gap ? qnan : rs;
        re = gap ? qnan : re;
      }
      rs = qmin(rs, s);
      re = qmax(re, e);
    }
  };
  // lanes that do not need the block (its range is inside their cached
  // interval) neither load nor use it
  auto process_block = [&](int c, int jb, int m, const double* hrow, double& rs, double& re, bool need) {
    Block bk = {};
    if (need) load_item(c, jb, bk);
    compute_block(bk, jb, need ? m : jb, hrow, rs, re);
  };
  const int W = htab_width(rint);
  if (g.s_lo) {
    // Source columns nearest first (d = 0, -1, +1, -2, ...), their blocks in
    // z order.  The near columns build most of the union; a farther column
    // or block whose range, widened by the largest dilation at that distance
    // H(d, 0), lies inside every lane's cached interval changes nothing (the
    // union only grows) and is neither loaded nor processed.
    //
    // Columns are checked in groups of gather_group(): a group's counts and
    // column ranges (first lo from the block summaries, last hi) are loaded
    // while the group before is processed, so the wavefront waits on memory
    // about once per group of skipped columns.  A column that is processed
    // has its count and first summary loaded again (one round trip), and the
    // next block's summary is loaded while the current block is processed.
    constexpr int G = gather_group(kLdsTable);
    const int nsteps = 2 * rint + 1;
    auto col_of_step = [&](int step) { return x + ((step & 1) ? -((step + 1) >> 1) : ((step + 1) >> 1)); };
    int pm[G];
    T plo[G], phi[G];
    auto fetch = [&](int grp) {
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const int step = grp * G + j;
        const int c = col_of_step(step);
        pm[j] = 0;
        plo[j] = phi[j] = T(0);
        if (step < nsteps && c >= 0 && c < g.nx) {  // uniform
          pm[j] = count_of(c);
          plo[j] = g.s_lo[summary_index(g.ny, g.caps, c, yc, 0)];
          phi[j] = g.c_hi[count_index(g.ny, c, yc)];
        }
      }
    };
    fetch(0);
    for (int grp = 0; grp * G < nsteps; ++grp) {
      unsigned need = 0;
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const int ad = (grp * G + j + 1) >> 1;
        const double hmax = ad <= rint ? htab[ad * W] : 0.0;
        const bool in = pm[j] == 0 || ok == 0 || (double(plo[j]) - hmax >= clo && double(phi[j]) + hmax <= chi);
        need |= ballot(!in) != 0 ? 1u << j : 0u;
      }
      if ((grp + 1) * G < nsteps) fetch(grp + 1);
      while (need != 0) {
        const int j = __builtin_ctz(need);
        need &= need - 1;
        const int step = grp * G + j;
        const int c = col_of_step(step);
        const int m = count_of(c);  // loaded with the first summary
        const int ad = (step + 1) >> 1;
        const double* hrow = htab + ad * W;
        const double hmax = hrow[0];
        double rs = qnan, re = qnan;
        // The checks mark the blocks each lane needs (bit k of `want`: block
        // jw + k); a round then processes every lane's next marked block, so
        // lanes needing different blocks of the column share a round.  Each
        // lane's blocks go in z order, so its run continues across them; the
        // marks of a window of 64 blocks see the cached interval from before
        // the window's blocks are processed (a block may be processed that
        // could have been skipped, none that is needed is skipped).
        uint64_t want = 0;
        int jb = 0, jw = 0;
        auto drain = [&]() {
          while (ballot(want != 0) != 0) {
            // a block parks at most B + 1 runs per lane: flush before one
            // that could fill a lane's list area
            if (ballot(sp > 0 && n + sp + B + 1 > ocap) != 0) flush_held();
            const bool has = want != 0;
            const int k = has ? __builtin_ctzll(want) : 0;
            want &= want - 1;
            process_block(c, jw + k * B, m, hrow, rs, re, has);
          }
        };
        // Two ways to mark a column's blocks.  Forward: every block in z
        // order, the next block's summary loaded while the current one is
        // checked.  Two-ended: a column's block ranges increase along it (its
        // triples are z-sorted and disjoint), so the blocks not inside the
        // cached interval are a prefix (lo below it) and a suffix (hi above
        // it); each lane scans from both ends to its first block inside, one
        // summary from each end per memory round trip, and the blocks between
        // are inside (fewer round trips; measured faster everywhere but for
        // FP32 storage at R <= 31, where its registers cost the fifth wave).
        constexpr bool kTwoEnded = !(kLdsTable && sizeof(T) == 4);
        const int nb = ok ? (m + B - 1) / B : 0;  // the lane's blocks in this column
        if (kTwoEnded && ballot(nb > 64) == 0) {
          int f = 0, bk = nb - 1;
          int fd = nb == 0, bd = nb == 0;
          while (ballot(fd == 0 || bd == 0) != 0) {
            T flo = T(0), fhi = T(0), klo = T(0), khi = T(0);
            if (fd == 0) {
              flo = g.s_lo[summary_index(g.ny, g.caps, c, yc, f * B)];
              fhi = g.s_hi[summary_index(g.ny, g.caps, c, yc, f * B)];
            }
            if (bd == 0) {
              klo = g.s_lo[summary_index(g.ny, g.caps, c, yc, bk * B)];
              khi = g.s_hi[summary_index(g.ny, g.caps, c, yc, bk * B)];
            }
            const bool fin = double(flo) - hmax >= clo && double(fhi) + hmax <= chi;
            const bool kin = double(klo) - hmax >= clo && double(khi) + hmax <= chi;
            want |= (fd == 0 && !fin) ? uint64_t(1) << f : 0;
            want |= (bd == 0 && !kin) ? uint64_t(1) << bk : 0;
            f += (fd == 0 && !fin) ? 1 : 0;
            bk -= (bd == 0 && !kin) ? 1 : 0;
            fd = (fd != 0 || fin || f > bk) ? 1 : 0;
            bd = (bd != 0 || kin || f > bk) ? 1 : 0;
          }
          drain();
        } else {
          T blo = g.s_lo[summary_index(g.ny, g.caps, c, yc, 0)];
          T bhi = g.s_hi[summary_index(g.ny, g.caps, c, yc, 0)];
          for (;;) {
            const bool more = ballot(m > jb + B && ok) != 0;
            T nblo = blo, nbhi = bhi;
            if (more) {
              nblo = g.s_lo[summary_index(g.ny, g.caps, c, yc, jb + B)];
              nbhi = g.s_hi[summary_index(g.ny, g.caps, c, yc, jb + B)];
            }
            const bool inside = jb >= m || ok == 0 || (double(blo) - hmax >= clo && double(bhi) + hmax <= chi);
            want |= inside ? 0 : uint64_t(1) << ((jb - jw) / B);
            if (!more || jb - jw == 63 * B) {
              drain();
              jw = jb + B;
            }
            if (!more) break;
            jb += B;
            blo = nblo;
            bhi = nbhi;
          }
        }
        unite(rs == rs && ok, rs, re);
        flush_held();
      }
    }
  } else {
    // Without summaries: blocks outermost, source columns nearest first
    // within a block, so a block's cache lines are read by all its steps
    // at about the same time.
    int jmax = 0;
    for (int c = c_lo; c <= c_hi; ++c) jmax = max(jmax, count_of(c));
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) jmax = max(jmax, __shfl_xor(jmax, d));
    for (int jb = 0; jb < jmax; jb += B) {
      for (int step = 0; step <= 2 * rint; ++step) {
        const int ad = (step + 1) >> 1;
        const int d = (step & 1) ? -ad : ad;
        const int c = x + d;
        if (c < 0 || c >= g.nx) continue;  // uniform
        const int m = count_of(c);
        if (ballot(m > jb && ok) == 0) continue;
        double rs = qnan, re = qnan;
        process_block(c, jb, m, htab + ad * W, rs, re, true);
        unite(rs == rs && ok, rs, re);
        flush_held();
      }
    }
  }
  flush_held();
  if (dirty && ok) {
    olo[k * kWave] = clo;
    ohi[k * kWave] = chi;
  }
  if (ballot(active && ok == 0) != 0) {
    if (!g.tile_list)
      raise_status(g.status, kStatusGatherOverflow);  // batched: the host redoes this volume alone
    else if (L == 0)
      g.tile_list[1 + atomicAdd(g.tile_list, 1)] = int(tile);
    return;
  }
  if (!active) return;
  // interval i of lane L at (tile base) + i * kOutStride + 2 L: the
  // wavefront's i-th intervals are one contiguous run
  const int64_t ob = g.o_base + tile * g.arena + 2 * int64_t(L);
  for (int i = 0; i < n; ++i) {
    g.o_z[ob + int64_t(i) * kOutStride] = T(olo[i * kWave]);
    g.o_z[ob + int64_t(i) * kOutStride + 1] = T(ohi[i * kWave]);
  }
  g.o_start[col] = ob;
  g.o_cnt[col] = 2 * n;
}

// Waves per SIMD the gather is compiled for: the kernel waits on memory most
// of the time, so occupancy pays until the registers spill: five for FP32
// storage at R <= 31 (at most 96 registers), four otherwise, where the
// prefetched column ranges, table reads, FP64 blocks and the two-ended block
// scan take more.
template <class T, bool kLdsTable>
struct GatherWaves {
  static constexpr int value = kLdsTable && sizeof(T) == 4 ? 5 : 4;
};

template <class T, bool kLdsTable>
__global__ __launch_bounds__(kWave, (GatherWaves<T, kLdsTable>::value)) void pass2_gather_kernel(Pass2GArgs<T> g) {
  gather_tile<T, kLdsTable>(g, xcd_tile(blockIdx.x, gridDim.x));
}

// Batched: job j owns the global tiles [tile_base[j], tile_base[j + 1]);
// this launch covers the n tiles from t0 on (one radius class), one per
// workgroup, or (kQueue) as a persistent grid taking them from the eight
// counters at `queue` (take_tiles).  The two forms are separate kernels: the
// tile body inlined in the queue loop needs registers the one-tile form
// would otherwise spill for.
template <class T, bool kLdsTable, bool kQueue>
__global__ __launch_bounds__(kWave, (GatherWaves<T, kLdsTable>::value)) void pass2_gather_batch_kernel(const Pass2GArgs<T>* jobs,
                                                                   const int64_t* tile_base, int njobs,
                                                                   int64_t t0, int64_t n, unsigned* queue) {
  auto body = [&](int64_t i) {
    const int64_t t = t0 + i;
    const int j = find_job(tile_base, njobs, t);
    const Pass2GArgs<T> g = jobs[j];
    gather_tile<T, kLdsTable>(g, t - tile_base[j]);
  };
  if constexpr (kQueue)
    take_tiles(queue, n, body);
  else
    body(xcd_chunk_tile(blockIdx.x, n));
}

}  // namespace vo

// ---------------------------------------------------------------------------
// Pass 1, painter form.  The same per-line sweep as pass1_kernel without the
// crossing-point culling: the live list holds, for every z, the nearest line
// (at most rint behind) whose input covers z, i.e. the newest piece wins.  A
// step subtracts the new column's segments from every live piece (lanes over
// pieces, segments broadcast from registers), places the surviving fragments
// and the new segments by prefix counts, and drops pieces older than rint.
// Culling only thins the intermediate (it never changes the dilation), and on
// Poisson volumes it removes under 15% of the triples, so this form skips it.
// ---------------------------------------------------------------------------

namespace vo {

__device__ __forceinline__ float lane_bcast(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double lane_bcast(double v, int j) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffff), j);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), j);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | (static_cast<unsigned int>(lo)));
}

template <class T, int CAP, int ARENA, int CMAX>
struct Pass1PSmem {
  static constexpr int ECAP = 4 * CAP;
  T lo[2][CAP];
  T hi[2][CAP];
  int key[2][CAP];
  int hist[kWave + 1];
  // forward results of the chunk
  T alo[ARENA];
  T ahi[ARENA];
  uint16_t ady[ARENA];
  int apos[CMAX];
  int acnt[CMAX];
  // union of the forward and backward results
  T ev[ECAP];
  int ec[ECAP];
  T tlo[ECAP];
  T thi[ECAP];
  int tcode[ECAP];
  int ka[CAP];
  int kb[CAP];
  T ulo[2 * CAP];
  T uhi[2 * CAP];
  int ukey[2 * CAP];
};

// Fragments of piece [lo, hi] (lo < hi) left by the batch's sorted disjoint
// segments (lane j holds segment j, j < nb).  Calls emit(flo, fhi, below) for
// each, where `below` is the number of batch segments under the fragment.
template <class T, class Emit>
__device__ __forceinline__ void piece_fragments(double lo, double hi, T nlo, T nhi, int nb, Emit&& emit) {
  double c = lo;
  bool done = false;
  for (int j = 0; j < nb; ++j) {
    const double a = lane_bcast(nlo, j), b = lane_bcast(nhi, j);
    if (done || b <= c || a >= hi) continue;
    if (a > c) emit(c, a, j);
    c = fmax(c, b);
    done = c >= hi;
  }
  if (!done) {
    int below = 0;
    for (int j = 0; j < nb; ++j) below += double(lane_bcast(nlo, j)) < hi ? 1 : 0;
    emit(c, hi, below);
  }
}

// A point piece survives unless a batch segment covers it; returns the number
// of batch segments below it, or -1 when covered.
template <class T>
__device__ __forceinline__ int point_below(double z, T nlo, T nhi, int nb) {
  int below = 0;
  bool covered = false;
  for (int j = 0; j < nb; ++j) {
    const double a = lane_bcast(nlo, j), b = lane_bcast(nhi, j);
    covered |= a <= z && z <= b;
    below += b < z ? 1 : 0;
  }
  return covered ? -1 : below;
}

// One painter step on line k; the column's segments are z[zoff .. zoff+2*np),
// the first batch of which is already in (pre_lo, pre_hi).  Returns the new
// size of list `cur` (which flips), or -1 on overflow.
template <class T, class S>
__device__ int paint_step(S& sm, int& cur, int n, int k, int rint, const T* z, int64_t zoff, int64_t np,
                          T pre_lo, T pre_hi, int cap, int* status) {
  const int L = lane_id();
  for (int64_t b0 = 0; b0 == 0 || b0 < np; b0 += kWave) {
    const int nb = int((np - b0) < kWave ? (np - b0) : int64_t(kWave));
    T nlo = pre_lo, nhi = pre_hi;
    if (b0 > 0 && L < nb) {
      nlo = z[zoff + 2 * (b0 + L)];
      nhi = z[zoff + 2 * (b0 + L) + 1];
    }
    if (L <= nb) sm.hist[L] = 0;
    wave_sync();
    const int c = cur;
    int placed = 0;  // fragments of earlier chunks
    for (int base = 0; base < n; base += kWave) {
      const int i = base + L;
      int f = 0;
      double plo = 0, phi = 0;
      int pkey = 0;
      bool live = false;
      if (i < n) {
        pkey = sm.key[c][i];
        live = k - pkey <= rint;
        plo = sm.lo[c][i];
        phi = sm.hi[c][i];
      }
      if (live) {
        if (plo < phi)
          piece_fragments(plo, phi, nlo, nhi, nb, [&](double, double, int) { ++f; });
        else
          f = point_below(plo, nlo, nhi, nb) >= 0 ? 1 : 0;
      }
      int v = f;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int o = __shfl_up(v, d);
        if (L >= d) v += o;
      }
      const int total = placed + __shfl(v, kWave - 1);
      if (total + nb > cap) {
        raise_status(status, kStatusP1List);
        return -1;
      }
      int q = placed + v - f;  // index of this piece's first fragment
      if (f > 0) {
        if (plo < phi) {
          piece_fragments(plo, phi, nlo, nhi, nb, [&](double flo, double fhi, int below) {
            const int pos = q + below;
            sm.lo[c ^ 1][pos] = T(flo);
            sm.hi[c ^ 1][pos] = T(fhi);
            sm.key[c ^ 1][pos] = pkey;
            atomicAdd(&sm.hist[below], 1);
            ++q;
          });
        } else {
          const int below = point_below(plo, nlo, nhi, nb);
          sm.lo[c ^ 1][q + below] = T(plo);
          sm.hi[c ^ 1][q + below] = T(phi);
          sm.key[c ^ 1][q + below] = pkey;
          atomicAdd(&sm.hist[below], 1);
        }
      }
      placed = total;
    }
    wave_sync();
    // batch segment j goes after the fragments with at most j segments below
    int h = L <= nb ? sm.hist[L] : 0;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int o = __shfl_up(h, d);
      if (L >= d) h += o;
    }
    if (L < nb) {
      sm.lo[c ^ 1][L + h] = nlo;
      sm.hi[c ^ 1][L + h] = nhi;
      sm.key[c ^ 1][L + h] = k;
    }
    n = placed + nb;
    cur ^= 1;
    wave_sync();
  }
  return n;
}

template <class T, int CAP, int ARENA, int CMAX>
__global__ __launch_bounds__(kWave) void pass1_paint_kernel(Pass1Args<T> g) {
  using S = Pass1PSmem<T, CAP, ARENA, CMAX>;
  __shared__ S sm;
  const int L = lane_id();
  const int x = blockIdx.x % g.nx;
  const int ci = blockIdx.x / g.nx;
  const int y0 = ci * g.chunk;
  const int y1 = min(g.ny, y0 + g.chunk);
  if (y0 >= g.ny) return;
  auto column = [&](int y) { return int64_t(x) * g.sx + int64_t(y) * g.sy; };
  // descriptor and first batch of a column, fetched one step ahead
  auto fetch = [&](int y, int64_t& zo, int64_t& np, T& a, T& b) {
    const int64_t col = column(y);
    zo = g.in_off[col];
    np = (g.in_off[col + 1] - zo) / 2;
    a = b = T(0);
    if (L < np) {
      a = g.in_z[zo + 2 * L];
      b = g.in_z[zo + 2 * L + 1];
    }
  };
  // forward sweep, k = y
  int cur = 0, n = 0, apos = 0;
  const int ys = max(0, y0 - g.rint);
  int64_t zo, np;
  T pa, pb;
  fetch(ys, zo, np, pa, pb);
  for (int y = ys; y < y1; ++y) {
    int64_t zo_n = 0, np_n = 0;
    T pa_n = T(0), pb_n = T(0);
    if (y + 1 < y1) fetch(y + 1, zo_n, np_n, pa_n, pb_n);
    n = paint_step<T>(sm, cur, n, y, g.rint, g.in_z, zo, np, pa, pb, CAP, g.status);
    if (n < 0) return;
    if (y >= y0) {
      if (apos + n > ARENA) {
        raise_status(g.status, kStatusArenaOverflow);
        return;
      }
      for (int i = L; i < n; i += kWave) {
        sm.alo[apos + i] = sm.lo[cur][i];
        sm.ahi[apos + i] = sm.hi[cur][i];
        sm.ady[apos + i] = uint16_t(y - sm.key[cur][i]);
      }
      if (L == 0) {
        sm.apos[y - y0] = apos;
        sm.acnt[y - y0] = n;
      }
      apos += n;
      wave_sync();
    }
    zo = zo_n;
    np = np_n;
    pa = pa_n;
    pb = pb_n;
  }
  // backward sweep, k = ny-1-y; unite with the forward result and store
  cur = 0;
  n = 0;
  const int ye = min(g.ny - 1, y1 - 1 + g.rint);
  fetch(ye, zo, np, pa, pb);
  for (int y = ye; y >= y0; --y) {
    const int k = g.ny - 1 - y;
    int64_t zo_n = 0, np_n = 0;
    T pa_n = T(0), pb_n = T(0);
    if (y - 1 >= y0) fetch(y - 1, zo_n, np_n, pa_n, pb_n);
    n = paint_step<T>(sm, cur, n, k, g.rint, g.in_z, zo, np, pa, pb, CAP, g.status);
    if (n < 0) return;
    zo = zo_n;
    np = np_n;
    pa = pa_n;
    pb = pb_n;
    if (y >= y1) continue;
    const int fa = sm.apos[y - y0], fn = sm.acnt[y - y0];
    for (int i = L; i < fn; i += kWave) sm.ka[i] = -int(sm.ady[fa + i]);
    for (int i = L; i < n; i += kWave) sm.kb[i] = -(k - sm.key[cur][i]);
    wave_sync();
    PieceRef<T> A{sm.alo + fa, sm.ahi + fa, sm.ka};
    PieceRef<T> B{sm.lo[cur], sm.hi[cur], sm.kb};
    PieceRef<T> U{sm.ulo, sm.uhi, sm.ukey};
    MergeScratch<T> ms{sm.ev, sm.ec, sm.tlo, sm.thi, sm.tcode, S::ECAP};
    const int nu = merge_max(A, fn, B, n, U, 2 * CAP, ms);
    if (nu < 0 || nu > g.caps) {
      raise_status(g.status, nu < 0 ? kStatusP1Union : kStatusSlotOverflow);
      return;
    }
    for (int i = L; i < nu; i += kWave) {
      const int64_t o = slot_index(g.ny, g.caps, x, y, i);
      g.m_lo[o] = sm.ulo[i];
      g.m_hi[o] = sm.uhi[i];
      g.m_dy[o] = uint16_t(-sm.ukey[i]);
    }
    if (L == 0) g.m_cnt[count_index(g.ny, x, y)] = nu;
  }
}

}  // namespace vo

// ---------------------------------------------------------------------------
// Pass 1, z-sweep form.  Without culling, the pass-1 result at column (x, y)
// is, for every z, the nearest line y' (|y - y'| <= R) whose input covers z,
// which is what the union of the forward and backward half sweeps of
// VoronoiVorPower.cpp:39-66 represents (both directions at once, nearer line
// winning).  One wavefront per (x-line, 64 output columns): the segment
// endpoints of the window of 64 + 2R input columns are counting-sorted by z
// in LDS, then swept in z order with a coverage bit per window column; every
// lane keeps the distance to its nearest covered column and emits a piece
// whenever that distance changes.  Pieces come out z-sorted, already united
// across both directions.  Windows with more endpoints than fit are swept in
// z-slabs of consecutive histogram bins.
// ---------------------------------------------------------------------------

namespace vo {

template <class T>
struct Pass1ZArgs {
  int nx, ny;
  int rint;
  const int64_t* in_off;
  const T* in_z;
  T* m_lo;  // slot layout, see Pass1Args
  T* m_hi;
  uint16_t* m_dy;
  int caps;
  int32_t* m_cnt;
  int* status;
  // block summaries at summary_index: lo of the block's first triple, hi of
  // its last one
  T* s_lo;
  T* s_hi;
  T* c_hi;  // per column, at count_index: hi of its last triple (when it has one)
  // CSR index of the column at (line x, position y) of this logical grid:
  // x * sx + y * sy (sx = 1, sy = nx sweeps along y, the reference's axis;
  // sx = ny, sy = 1 with nx, ny swapped sweeps along x)
  int64_t sx, sy;
  // The input check of the operation, done here instead of by its own
  // kernel (null: none): each tile checks its 64 output columns' input as
  // the window is read, with the rules and words of check_input_kernel.
  int* check_word = nullptr;
  int check_range = 0;
  double check_lo = 0, check_hi = 0;
};

constexpr int kZTile = kWave;    // output columns per wavefront
constexpr int kWalkMaxR = 40;    // largest R whose lanes walk their own events

// The z-sweep stores a wavefront's triples and summaries at 32-bit byte
// offsets from the wavefront's first slot and summary (a scalar base plus a
// vector offset); true when every offset fits.  The largest slot offset is
// below caps * ny + 512.
__host__ __device__ constexpr bool zsweep_offsets_fit(int caps, int ny, int tsize) {
  return (int64_t(caps) * ny + 1024) * tsize < (int64_t(1) << 32);
}
constexpr int kMaxSlabs = 64;    // z-slabs per window

// Window columns for the lanes' event masks (walking lanes, R <= kWalkMaxR):
// bit i of colmask[w] marks event i of the current chunk as one of window
// column w's.
constexpr int kColMaskMax = kWave + 2 * kWalkMaxR;  // (also >= kWave + 1: the cell walk's difference array)

template <class T, int ECAP, int NB, int WMAX>
struct Pass1ZSmem {
  T ez[ECAP];
  uint16_t ep[ECAP];  // (window column << 2) | kind: 0 = start, 1 = point, 2 = end
  static_assert((WMAX << 2) < 65536, "window columns fit the event code");
  int bin[NB + 1];
  int slab_end[kMaxSlabs + 1];
  uint64_t colmask[kColMaskMax];
};

// Visits the merged segments of one input column (touching or overlapping
// neighbours fused, points kept only when isolated).
// Two segments are loaded per step, so that a column costs half as many
// memory round trips.
// `raw` sees every stored pair (a, b) in order before it is merged.
template <class T, class F, class P>
__device__ __forceinline__ void column_segments(const T* z, int64_t s, int64_t n, F&& f, P&& raw) {
  if (n < 2) return;
  T ca = z[s], cb = z[s + 1];
  raw(ca, cb);
  auto step = [&](T a, T b) {
    raw(a, b);
    if (a <= cb) {
      if (b > cb) cb = b;
    } else {
      f(ca, cb);
      ca = a;
      cb = b;
    }
  };
  int64_t k = s + 2;
  for (; k + 3 < s + n; k += 4) {
    const T a0 = z[k], b0 = z[k + 1], a1 = z[k + 2], b1 = z[k + 3];
    step(a0, b0);
    step(a1, b1);
  }
  if (k + 1 < s + n) step(z[k], z[k + 1]);
  f(ca, cb);
}
template <class T, class F>
__device__ __forceinline__ void column_segments(const T* z, int64_t s, int64_t n, F&& f) {
  column_segments(z, s, n, f, [](T, T) {});
}

// base[i] as a scalar base plus a 32-bit vector byte offset
template <class V>
__device__ __forceinline__ void store_at(V* base, uint32_t i, V v) {
  *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + i * uint32_t(sizeof(V))) = v;
}

// Where a lane's pass-1 triples go.  With kRegs (FP32 storage, R <= 31) the
// lane's current block of kSlotBlock triples is kept in registers (a shift
// register: the newest triple enters at the top) and written with 16-byte
// stores when it fills, together with its summary, so every store covers
// whole 32-byte sectors of the slot layout and the summaries need no second
// pass.  Otherwise (no registers to spare beside the wider lane state) each
// triple is stored as it comes and the summaries are written at the end from
// the stored blocks.
template <class T, bool kRegs>
struct BlockStage {
  static constexpr int B = kSlotBlock;
  // the wavefront's first slot in each array and its first summary (uniform)
  T* lo_w;
  T* hi_w;
  uint16_t* dy_w;
  T* slo_w;
  T* shi_w;
  uint32_t L, ny;
  int caps;
  // slots 0-3 and 4-7 as 4-vectors, in the registers of the 16-byte stores
  float4 la = {}, lb = {}, ha = {}, hb = {};
  uint64_t dy = 0;  // the block's line offsets, 8 bits each (offsets <= 127)

  __device__ __forceinline__ uint32_t slot(int j) const {
    return L * B + uint32_t(j / B) * (ny * B) + uint32_t(j % B);
  }
  // writes stage block `blk` (its triples in slots 0 .. B - 1) and its
  // summary (lo of the first triple, hi of the last)
  __device__ __forceinline__ void store_block(int blk, T last) {
    const uint32_t s = slot(blk * B);
    float4* lw = reinterpret_cast<float4*>(reinterpret_cast<char*>(lo_w) + s * 4u);
    float4* hw = reinterpret_cast<float4*>(reinterpret_cast<char*>(hi_w) + s * 4u);
    lw[0] = la;
    lw[1] = lb;
    hw[0] = ha;
    hw[1] = hb;
    // 8-bit offsets widened to the 16-bit slots: bytes (b0, 0, b1, 0), ...
    const uint32_t d0 = uint32_t(dy), d1 = uint32_t(dy >> 32);
    const uint4 q = make_uint4(__builtin_amdgcn_perm(0u, d0, 0x0c010c00u), __builtin_amdgcn_perm(0u, d0, 0x0c030c02u),
                               __builtin_amdgcn_perm(0u, d1, 0x0c010c00u), __builtin_amdgcn_perm(0u, d1, 0x0c030c02u));
    *reinterpret_cast<uint4*>(reinterpret_cast<char*>(dy_w) + s * 2u) = q;
    store_at(slo_w, L + uint32_t(blk) * ny, T(la.x));
    store_at(shi_w, L + uint32_t(blk) * ny, last);
  }
  // one step of the shift register: slot i takes slot i + 1, slot 7 takes v
  __device__ __forceinline__ static void shift_in(bool on, float4& a, float4& b, float v) {
    a.x = on ? a.y : a.x;
    a.y = on ? a.z : a.y;
    a.z = on ? a.w : a.z;
    a.w = on ? b.x : a.w;
    b.x = on ? b.y : b.x;
    b.y = on ? b.z : b.y;
    b.z = on ? b.w : b.z;
    b.w = on ? v : b.w;
  }
  // triple j = (a, b, d) of the lane where `put`
  __device__ __forceinline__ void put(bool on, int j, T a, T b, int d) {
    if constexpr (kRegs) {
      shift_in(on, la, lb, float(a));
      shift_in(on, ha, hb, float(b));
      dy = on ? (dy >> 8) | (uint64_t(uint32_t(d)) << 56) : dy;
      if (on && j % B == B - 1 && j < caps) store_block(j / B, b);
    } else {
      if (on && j < caps) {
        const uint32_t s = slot(j);
        store_at(lo_w, s, a);
        store_at(hi_w, s, b);
        store_at(dy_w, s, uint16_t(d));
      }
    }
  }
  // after the lane's last triple (count triples, the last with hi `last`)
  __device__ __forceinline__ void finish(int count, T last, bool active) {
    const int kept = min(count, caps);
    if constexpr (kRegs) {
      // a partial last block: its n triples sit at the top of the stage;
      // shift them down by B - n into slots 0 .. n - 1
      const int n = count % B;
      const bool part = active && n > 0 && count <= caps;
      const int sh = B - n;
      for (int k = 0; k < B - 1; ++k) {  // at most 7 single steps, same for all lanes
        const bool m = part && k < sh;
        if (ballot(m) == 0) break;
        shift_in(m, la, lb, 0.f);
        shift_in(m, ha, hb, 0.f);
        dy = m ? dy >> 8 : dy;
      }
      if (part) store_block(count / B, last);
    } else {
      // summaries from the stored blocks, block by block for all lanes
      const int nblk = active ? (kept + B - 1) / B : 0;
      const int nblk_max = wave_max(nblk);
      for (int b = 0; b < nblk_max; ++b) {
        if (b < nblk) {
          const uint32_t s = slot(b * B);
          const int lastj = min(kept - b * B, B) - 1;
          store_at(slo_w, L + uint32_t(b) * ny, lo_w[s]);
          store_at(shi_w, L + uint32_t(b) * ny, hi_w[s + uint32_t(lastj)]);
        }
      }
    }
  }
};

// Sweep state of one lane.  The lane's own column is window position L + R;
// a covered window column L + k within reach (0 <= k <= 2R) is kept as one
// bit indexed by its distance d = |k - R|, so the distance to the nearest
// covered column is one bit scan.  With NW = 1 (R <= 31) the sides
// interleave in one 64-bit word: bit 2d for k >= R, bit 2d + 1 for k < R, and
// the distance is half the lowest set bit's index.  With NW = 2 or 4 each side
// has NW / 2 words of its own (bit d of the side's words; R < 32 NW) and the
// distance is the lowest set bit of their union.  Within one column the
// merged segments neither overlap nor touch, so every start or end event
// flips its column's bit.
// The sweep only compares endpoint values it read, so the open piece's start
// is kept in the storage type T.
template <int NW, class T>
struct ZState {
  uint64_t w[NW];
  int R = 0;
  int dist = 1 << 20, dcur = 1 << 20, count = 0;
  T z0 = T(0);
  // loop-carried flags as 0/1 integers: in vector registers they need no
  // exec-mask merges at every join of the walk
  int point_open = 0, ok = 1;
  T last_hi = T(0);  // hi of the lane's last triple
  // the lane's pending group of points of equal z (their bits stay set until
  // the lane's next event that is not a point at the same z)
  int pend = 0;
  T pz = T(0);
  uint64_t gpts[NW];

  __device__ __forceinline__ explicit ZState(int r) : R(r) {
#pragma unroll
    for (int t = 0; t < NW; ++t) w[t] = gpts[t] = 0;
  }
  static constexpr int H = NW == 1 ? 1 : NW / 2;  // words per side (NW > 1)
  // distance to the nearest covered column, or R + 1 when none is
  __device__ __forceinline__ int nearest() const {
    if constexpr (NW == 1) {
      return w[0] ? int(__builtin_ctzll(w[0])) >> 1 : R + 1;
    } else {
      int best = R + 1;
#pragma unroll
      for (int t = H - 1; t >= 0; --t) {
        const uint64_t m = w[t] | w[H + t];
        best = m ? 64 * t + int(__builtin_ctzll(m)) : best;
      }
      return best;
    }
  }
  // the bits of window column k (0 <= k <= 2R) where `on`, word by word
  __device__ __forceinline__ void column_bits(int k, bool on, uint64_t (&bits)[NW]) const {
    if constexpr (NW == 1) {
      const int idx = k >= R ? 2 * (k - R) : 2 * (R - k) + 1;
      bits[0] = on ? (uint64_t(1) << idx) : 0;
    } else {
      const bool upper = k >= R;
      const int d = upper ? k - R : R - k;
      const uint64_t b = on ? (uint64_t(1) << (d & 63)) : 0;
#pragma unroll
      for (int t = 0; t < H; ++t) {
        const uint64_t bt = (H == 1 || t == (d >> 6)) ? b : 0;
        bits[t] = upper ? bt : 0;
        bits[H + t] = upper ? 0 : bt;
      }
    }
  }
  // The updates below are written with selects, predicated by `on`: lanes
  // walk their own events in lock step, and only the store of a finished
  // piece is a predicated branch.
  //
  // Starts and ends are applied one at a time: within a group of equal
  // (z, kind) the distance moves one way only, so the pieces a group would
  // close at once are closed at its first change and the pieces in between
  // have zero length and are not kept.  Points are grouped: all points of
  // one z are in place together for one zero-length piece.

  // close the open piece and open the next one when the distance changed;
  // zero-length pieces are kept only when a point opened them
  template <class S>
  __device__ __forceinline__ void emit(T z, bool is_point, bool on, bool active, const Pass1ZArgs<T>& g, S& stage) {
    const bool changed = on && dist != dcur;
    const bool put = changed && dcur <= R && (z0 < z || point_open != 0) && active;
    stage.put(put, count, z0, z, dcur);
    ok = (put && count >= g.caps) ? 0 : ok;
    last_hi = put ? z : last_hi;
    count += put ? 1 : 0;
    const bool open = changed && dist <= R;
    z0 = open ? z : z0;
    point_open = open ? int(is_point) : point_open;
    dcur = on ? dist : dcur;
  }
  // a finished point group: pieces are closed at its z with the points in
  // place, the points withdrawn and the piece after them opened at that z
  template <class S>
  __device__ __forceinline__ void close_points(bool on, bool active, const Pass1ZArgs<T>& g, S& stage) {
    emit(pz, true, on, active, g, stage);
#pragma unroll
    for (int t = 0; t < NW; ++t) {
      w[t] &= on ? ~gpts[t] : ~uint64_t(0);
      gpts[t] = on ? 0 : gpts[t];
    }
    dist = nearest();
    emit(pz, false, on, active, g, stage);
    pend = on ? 0 : pend;
  }
  // one event of window column L + k within reach (0 <= k <= 2R) where `on`;
  // kind 0 start, 1 point, 2 end
  template <class S>
  __device__ __forceinline__ void apply(T z, int kind, int k, bool on, bool active, const Pass1ZArgs<T>& g,
                                       S& stage) {
    const bool point = kind == 1;
    const bool flush = on && pend != 0 && !(point && z == pz);
    if (ballot(flush) != 0) close_points(flush, active, g, stage);
    uint64_t bits[NW];
    column_bits(k, on, bits);
#pragma unroll
    for (int t = 0; t < NW; ++t) {
      w[t] ^= bits[t];
      gpts[t] |= point ? bits[t] : 0;
    }
    dist = nearest();
    pz = (on && point) ? z : pz;
    pend = (on && point) ? 1 : pend;
    emit(z, false, on && !point, active, g, stage);
  }
};

// Pass 1 as a walk over Voronoi cells (windows without zero-length
// segments).  At height z the covered window columns split the line into
// cells: output column L (window position L + R) is in the cell of its
// nearest covered column, and its distance d_L is what the triples record.
// An event at column c changes d only inside c's cell: for a start, the
// positions strictly nearer to c than to the covered columns p < c and
// q > c around it (d becomes |c - pos|); for an end, the positions whose
// nearest was c alone (d becomes the distance to p or q, or "none" beyond
// R).  So each event touches a contiguous run of a few output columns, not
// all 2R + 1 within reach, and the walk costs about one step per triple.
//
// Events are read 64 at a time in (z, kind) order.  Lane i holds event i:
// an inclusive XOR scan of one-hot column bits over the lanes, on top of the
// covered mask before the chunk, gives the mask before each event; p and q
// are bit scans in it; the event's run of output columns [a, b] becomes
// bit i of the lanes in the run through an XOR difference array (the bits
// are distinct, so OR = XOR) and a prefix XOR.  Then every lane walks its
// own events in order, as the lane walk does.
constexpr int kCellMaxR = 255;

template <class T, int MW>
struct CellWalk {
  // covered window columns before the current chunk (uniform)
  uint64_t carry[MW];
  int R;
  bool active;
  // the lane's state: its open piece starts at z0 with distance dcur
  // (R + 1: no piece)
  T z0 = T(0);
  T last_hi = T(0);
  int dcur, count = 0, ok = 1;

  __device__ __forceinline__ CellWalk(int r, bool act) : R(r), active(act), dcur(r + 1) {
#pragma unroll
    for (int t = 0; t < MW; ++t) carry[t] = 0;
  }

  // an event word: c, p + 1, q + 1 in 10 bits each (window columns < 1023), start flag above
  static constexpr int kBits = 10, kNone = (1 << kBits) - 1;

  // sweep the E events (ez, ep) of one z-slab; `diff` is LDS scratch of
  // kWave + 1 words
  template <class A, class S>
  __device__ void sweep(const T* ez, const uint16_t* ep, int E, uint64_t* diff, const A& g, S& stage) {
    const int L = lane_id();
    for (int base = 0; base < E; base += kWave) {
      const int cnt = min(kWave, E - base);
      const bool has = L < cnt;
      const T vz = has ? ez[base + L] : T(0);
      const int code = has ? ep[base + L] : 0;
      const int c = code >> 2;
      const bool start = (code & 3) == 0;
      // one-hot column bits, inclusive XOR scan over the lanes
      uint64_t x[MW], own[MW];
#pragma unroll
      for (int t = 0; t < MW; ++t) {
        own[t] = (has && (c >> 6) == t) ? (uint64_t(1) << (c & 63)) : 0;
        x[t] = own[t];
      }
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
#pragma unroll
        for (int t = 0; t < MW; ++t) {
          const uint64_t o = __shfl_up(x[t], d);
          x[t] ^= L >= d ? o : 0;
        }
      }
      // covered columns other than c, just before (start) or after (end)
      // the event: the mask before it without c
      uint64_t m[MW];
#pragma unroll
      for (int t = 0; t < MW; ++t) m[t] = (carry[t] ^ x[t] ^ own[t]) & ~own[t];
#pragma unroll
      for (int t = 0; t < MW; ++t) carry[t] ^= __shfl(x[t], cnt - 1);
      // p: highest covered column below c, q: lowest above (kNone: none)
      int p = -1, q = 1 << 20;
#pragma unroll
      for (int t = 0; t < MW; ++t) {
        const int lo_c = c - 64 * t;  // bits of word t below c: [0, lo_c)
        const uint64_t below = lo_c <= 0 ? 0 : (lo_c >= 64 ? m[t] : (m[t] & ((uint64_t(1) << lo_c) - 1)));
        p = below ? 64 * t + 63 - __builtin_clzll(below) : p;
      }
#pragma unroll
      for (int t = MW - 1; t >= 0; --t) {
        const int hi_c = c + 1 - 64 * t;  // bits of word t above c: [hi_c, 64)
        const uint64_t above = hi_c >= 64 ? 0 : (hi_c <= 0 ? m[t] : (m[t] & (~uint64_t(0) << hi_c)));
        q = above ? 64 * t + __builtin_ctzll(above) : q;
      }
      // c's cell among the output positions [R, R + 63], within reach of c
      int a = max(c - R, p >= 0 ? ((p + c) >> 1) + 1 : c - R);
      int b = min(c + R, q < (1 << 20) ? ((c + q + 1) >> 1) - 1 : c + R);
      a = max(a, R) - R;
      b = min(b, R + kWave - 1) - R;
      // bit L of the run [a, b]: XOR difference array, then a prefix XOR
      if (L <= kWave) diff[L] = 0;
      wave_sync();
      if (has && a <= b) {
        const unsigned long long bit = 1ull << L;
        atomicXor(reinterpret_cast<unsigned long long*>(&diff[a]), bit);
        atomicXor(reinterpret_cast<unsigned long long*>(&diff[b + 1]), bit);
      }
      wave_sync();
      uint64_t mine = diff[L];
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const uint64_t o = __shfl_up(mine, d);
        mine ^= L >= d ? o : 0;
      }
      // the event as one word for the lanes that read it: c, p + 1, q + 1
      // (kNone beyond the window), start flag
      const int pk = c | ((p >= 0 ? p + 1 : 0) << kBits) | ((q < (1 << 20) ? q + 1 : kNone) << (2 * kBits)) |
                     (start ? 1 << (3 * kBits) : 0);
      const int pos = L + R;
      for (;;) {
        const bool on = mine != 0;
        if (ballot(on) == 0) break;
        const int i = on ? int(__builtin_ctzll(mine)) : 0;
        mine &= mine - 1;
        const T z = __shfl(vz, i);
        const int e = __shfl(pk, i);
        const int ec = e & kNone, ep1 = (e >> kBits) & kNone, eq1 = (e >> (2 * kBits)) & kNone;
        int dn;
        if (e >> (3 * kBits)) {
          dn = abs(ec - pos);
        } else {
          const int dp = ep1 ? pos - (ep1 - 1) : 1 << 20;
          const int dq = eq1 != kNone ? (eq1 - 1) - pos : 1 << 20;
          dn = min(min(dp, dq), R + 1);
        }
        const bool put = on && dcur <= R && z0 < z && active;
        stage.put(put, count, z0, z, dcur);
        ok = (put && count >= g.caps) ? 0 : ok;
        last_hi = put ? z : last_hi;
        count += put ? 1 : 0;
        z0 = on ? z : z0;
        dcur = on ? dn : dcur;
      }
    }
  }
};

// Bit i of the result: event i of the chunk (held by lane i, window column
// `col`) lies within this lane's reach, col - L in [0, 2R].  colmask[w]
// collects the chunk's events of window column w (one LDS atomic per event);
// the lane's mask is the OR of colmask over [L, L + 2R], built by doubling in
// registers: after the step with shift s every entry w covers [w, w + 2s).
// Entries w = L, L + 64, L + 128 live in lane L (wn <= kColMaskMax).
__device__ __forceinline__ uint64_t lane_event_mask(uint64_t* colmask, int col, bool has, int wn, int R) {
  const int L = lane_id();
  for (int w = L; w < wn; w += kWave) colmask[w] = 0;
  wave_sync();
  if (has) atomicOr(reinterpret_cast<unsigned long long*>(&colmask[col]), 1ull << L);
  wave_sync();
  const bool three = wn > 2 * kWave;  // uniform
  uint64_t e0 = colmask[L];
  uint64_t e1 = L + kWave < wn ? colmask[L + kWave] : 0;
  uint64_t e2 = three && L + 2 * kWave < wn ? colmask[L + 2 * kWave] : 0;
  const int W = 2 * R + 1;
  int span = 1;
  for (; 2 * span <= W; span *= 2) {
    const int src = (L + span) & (kWave - 1);
    const bool wrap = L + span >= kWave;
    const uint64_t a0 = __shfl(e0, src), a1 = __shfl(e1, src);
    uint64_t a2 = 0;
    if (three) a2 = __shfl(e2, src);
    e0 |= wrap ? a1 : a0;
    e1 |= wrap ? a2 : a1;
    e2 |= wrap ? 0 : a2;
  }
  // [L, L + W) = [L, L + span) | [L + W - span, L + W), with W - span < span <= 64
  const int off = W - span;
  const int src = (L + off) & (kWave - 1);
  const uint64_t b0 = __shfl(e0, src), b1 = __shfl(e1, src);
  return e0 | (L + off >= kWave ? b1 : b0);
}

template <class T, int ECAP, int NB, int WMAX, int NW>
__device__ void zsweep_tile(const Pass1ZArgs<T>& g, int64_t tile, Pass1ZSmem<T, ECAP, NB, WMAX>& sm) {
  static_assert(NB <= 32 * kWave, "one bit per lane's bin chunk in a 32-bit mask");
  const int L = lane_id();
  // tiles of one line are consecutive: neighbouring tiles share 2R columns
  // of their windows, and xcd_tile hands them to one XCD, whose L2 then
  // serves the overlap (and the window's second and third reads)
  const int64_t nyt = (g.ny + kZTile - 1) / kZTile;
  const int x = int(tile / nyt);
  const int y0 = int(tile % nyt) * kZTile;
  if (y0 >= g.ny) return;
  const int R = g.rint;
  const int wlo = y0 - R, wn = kZTile + 2 * R;
  auto col_range = [&](int w, int64_t& s, int64_t& n) {
    const int yy = wlo + w;
    s = n = 0;
    if (yy < 0 || yy >= g.ny) return;
    const int64_t c = int64_t(x) * g.sx + int64_t(yy) * g.sy;
    s = g.in_off[c];
    n = g.in_off[c + 1] - s;
  };
  // the input check of the tile's own columns (window positions R .. R + 63):
  // offsets here, endpoints as the histogram pass reads them
  const bool chk = g.check_word != nullptr;
  int bad = 0;
  if (chk && y0 + L < g.ny) {
    const int64_t c = int64_t(x) * g.sx + int64_t(y0 + L) * g.sy;
    const int64_t s = g.in_off[c], e = g.in_off[c + 1];
    bad = (e < s || ((e - s) & 1)) ? kInputMalformed : 0;
  }
  auto raw_check = [&](bool own, T& prev) {
    return [&, own](T a, T b) {
      // in order and finite (every comparison with a NaN fails)
      const bool okv = a >= prev && b >= a && fabs(a) < T(INFINITY) && fabs(b) < T(INFINITY);
      int m = own && !okv ? kInputMalformed : 0;
      if (g.check_range)  // uniform
        m = m ? m : ((own && (double(a) < g.check_lo || double(b) > g.check_hi)) ? kInputOutOfRange : 0);
      bad |= m;
      prev = b;
    };
  };
  auto report_check = [&]() {
    if (chk && bad) {
      atomicOr(g.check_word, bad);
      atomicMin(g.check_word + 1, 0);
    }
  };
  // z-range of the window
  double zlo = INFINITY, zhi = -INFINITY;
  for (int w = L; w < wn; w += kWave) {
    int64_t s, n;
    col_range(w, s, n);
    if (n >= 2) {
      zlo = fmin(zlo, double(g.in_z[s]));
      zhi = fmax(zhi, double(g.in_z[s + n - 1]));
    }
  }
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    zlo = fmin(zlo, __shfl_xor(zlo, d));
    zhi = fmax(zhi, __shfl_xor(zhi, d));
  }
  const bool active = y0 + L < g.ny;
  if (!(zlo <= zhi)) {  // no input anywhere in the window (or none that is finite)
    if (active) g.m_cnt[count_index(g.ny, x, y0 + L)] = 0;
    if (chk && active) {
      int64_t s, n;
      col_range(L + R, s, n);
      T prev = T(-INFINITY);
      for (int64_t i = 0; i + 1 < n; i += 2) raw_check(true, prev)(g.in_z[s + i], g.in_z[s + i + 1]);
    }
    report_check();
    return;
  }
  const double scale = double(NB) / fmax(zhi - zlo, 1e-30) * (1.0 - 1e-12);
  auto bin_of = [&](T z) { return min(NB - 1, max(0, int((double(z) - zlo) * scale))); };
  // histogram of the window's endpoints
  for (int b = L; b <= NB; b += kWave) sm.bin[b] = 0;
  wave_sync();
  int points = 0;
  for (int w = L; w < wn; w += kWave) {
    int64_t s, n;
    col_range(w, s, n);
    T prev = T(-INFINITY);
    column_segments(
        g.in_z, s, n,
        [&](T a, T b) {
          atomicAdd(&sm.bin[bin_of(a)], 1);
          if (a < b) atomicAdd(&sm.bin[bin_of(b)], 1);
          points |= a < b ? 0 : 1;
        },
        raw_check(chk && w >= R && w < R + kZTile && y0 + (w - R) < g.ny, prev));
  }
  report_check();
  // windows without zero-length segments take the cell walk (R <= 255: its
  // covered mask has 2, 3, 5 or 9 words for the window's 64 + 2R columns)
  const bool cells = ballot(points != 0) == 0 && R <= kCellMaxR;
  wave_sync();
  // slabs of consecutive bins holding at most ECAP endpoints each: one slab
  // when everything fits (the usual case), else a greedy split
  int total = 0;
  for (int b = L; b < NB; b += kWave) total += sm.bin[b];
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) total += __shfl_xor(total, d);
  int maxbin = 0;
  for (int b = L; b < NB; b += kWave) maxbin = max(maxbin, sm.bin[b]);
  maxbin = wave_max(maxbin);
  // Split by prefix: with S = ECAP - (largest bin) > (largest bin), slab k
  // holds the bins whose exclusive prefix lies in [k S, (k + 1) S), at most
  // S - 1 + (largest bin) < ECAP endpoints, and consecutive bins are in the
  // same or the next slab; the boundaries come from one scan of the bins.
  const int S = ECAP - maxbin;
  const bool by_prefix = total > ECAP && S > maxbin && (total - 1) / S + 1 <= kMaxSlabs;
  if (total <= ECAP) {
    if (L == 0) {
      sm.slab_end[0] = NB;
      sm.slab_end[kMaxSlabs] = 1;
    }
  } else if (by_prefix) {
    // trailing empty bins (exclusive prefix == total) join the last slab
    const int last_slab = (total - 1) / S;
    int carry = 0, prev_last = 0;
    for (int base = 0; base < NB; base += kWave) {
      const int b = base + L;
      const int c = b < NB ? sm.bin[b] : 0;
      int v = c;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int o = __shfl_up(v, d);
        if (L >= d) v += o;
      }
      const int slab = min((carry + v - c) / S, last_slab);
      int prev = __shfl_up(slab, 1);
      prev = L == 0 ? prev_last : prev;
      if (b < NB && b > 0 && slab != prev) sm.slab_end[prev] = b;
      prev_last = __shfl(slab, kWave - 1);
      carry += __shfl(v, kWave - 1);
    }
    if (L == 0) {
      sm.slab_end[prev_last] = NB;
      sm.slab_end[kMaxSlabs] = prev_last + 1;
    }
  } else if (L == 0) {  // greedy, one bin at a time
    int nslab = 0, acc = 0;
    for (int b = 0; b < NB && nslab >= 0; ++b) {
      const int c = sm.bin[b];
      if (c > ECAP) {
        nslab = -1;
        break;
      }
      if (acc + c > ECAP) {
        if (nslab + 1 >= kMaxSlabs) {
          nslab = -1;
          break;
        }
        sm.slab_end[nslab++] = b;
        acc = 0;
      }
      acc += c;
    }
    if (nslab >= 0) sm.slab_end[nslab++] = NB;
    sm.slab_end[kMaxSlabs] = nslab;
  }
  wave_sync();
  const int nslab = sm.slab_end[kMaxSlabs];
  if (nslab < 0) {
    raise_status(g.status, kStatusP1List);
    return;
  }
  // per-lane sweep state (lane = output column y0 + L, window position L + R)
  if (NW == 1 ? 2 * R + 1 >= 64 : R >= 32 * NW) {  // the host picks NW from R; never taken
    raise_status(g.status, kStatusP1List);
    return;
  }
  ZState<NW, T> st(R);
  const int64_t wslot = slot_index(g.ny, g.caps, x, y0, 0);
  const int64_t wsum = summary_index(g.ny, g.caps, x, y0, 0);
  BlockStage<T, sizeof(T) == 4 && NW == 1> stage;
  stage.lo_w = g.m_lo + wslot;
  stage.hi_w = g.m_hi + wslot;
  stage.dy_w = g.m_dy + wslot;
  stage.slo_w = g.s_lo + wsum;
  stage.shi_w = g.s_hi + wsum;
  stage.L = uint32_t(L);
  stage.ny = uint32_t(g.ny);
  stage.caps = g.caps;
  // covered-mask words of the cell walk: the window's 64 + 2R columns
  CellWalk<T, NW == 1 ? 2 : (NW == 2 ? 3 : (NW == 4 ? 5 : 9))> cw(R, active);
  int b_lo = 0;
  for (int sl = 0; sl < nslab; ++sl) {
    const int b_hi = sm.slab_end[sl];
    wave_sync();
    // exclusive bin starts of this slab; bit t of `unsorted`: the lane's bin
    // b_lo + L + 64 t holds two events or more
    uint32_t unsorted = 0;
    {
      int carry = 0;
      for (int base = b_lo; base < b_hi; base += kWave) {
        const int b = base + L;
        const int c = b < b_hi ? sm.bin[b] : 0;
        unsorted |= c >= 2 ? 1u << ((base - b_lo) / kWave) : 0u;
        int v = c;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
          const int o = __shfl_up(v, d);
          if (L >= d) v += o;
        }
        if (b < b_hi) sm.bin[b] = carry + v - c;
        carry += __shfl(v, kWave - 1);
      }
      wave_sync();
      if (L == 0) sm.bin[NB] = carry;
    }
    wave_sync();
    const int E = sm.bin[NB];
    // scatter this slab's endpoints (kind: 0 start, 1 point, 2 end)
    for (int w = L; w < wn; w += kWave) {
      int64_t s, n;
      col_range(w, s, n);
      column_segments(g.in_z, s, n, [&](T a, T b) {
        const int ba = bin_of(a);
        if (!(a < b)) {
          if (ba >= b_lo && ba < b_hi) {
            const int q = atomicAdd(&sm.bin[ba], 1);
            sm.ez[q] = a;
            sm.ep[q] = uint16_t((w << 2) | 1);
          }
          return;
        }
        if (ba >= b_lo && ba < b_hi) {
          const int q = atomicAdd(&sm.bin[ba], 1);
          sm.ez[q] = a;
          sm.ep[q] = uint16_t((w << 2) | 0);
        }
        const int bb = bin_of(b);
        if (bb >= b_lo && bb < b_hi) {
          const int q = atomicAdd(&sm.bin[bb], 1);
          sm.ez[q] = b;
          sm.ep[q] = uint16_t((w << 2) | 2);
        }
      });
    }
    wave_sync();
    // every bin now holds its end offset: order each bin of two events or
    // more by (z, kind)
    for (; unsorted != 0; unsorted &= unsorted - 1) {
      const int b = b_lo + L + kWave * int(__builtin_ctz(unsorted));
      const int e0 = (b == b_lo) ? 0 : sm.bin[b - 1];
      const int e1 = sm.bin[b];
      for (int i = e0 + 1; i < e1; ++i) {
        const T zi = sm.ez[i];
        const int pi = sm.ep[i];
        int j = i - 1;
        while (j >= e0 && (sm.ez[j] > zi || (sm.ez[j] == zi && (sm.ep[j] & 3) > (pi & 3)))) {
          sm.ez[j + 1] = sm.ez[j];
          sm.ep[j + 1] = sm.ep[j];
          --j;
        }
        sm.ez[j + 1] = zi;
        sm.ep[j + 1] = uint16_t(pi);
      }
    }
    wave_sync();
    // Sweep in (z, kind) order: the cell walk when it applies, else the lane
    // walk below.
    if (cells) {
      cw.sweep(sm.ez, sm.ep, E, sm.colmask, g, stage);
      b_lo = b_hi;
      continue;
    }
    // Lane walk: events are read 64 at a time into registers; a lane's
    // events are those of the 2R + 1 window columns within its reach
    // (k = column - L in [0, 2R]).  For R <= kWalkMaxR each lane marks its
    // events of the chunk and, when no lane has more than 3/4 of the chunk
    // marked, the lanes walk their own marked events in lock step, fetching
    // each with a lane permute.  Otherwise (and always for larger R, where a
    // lane's share of the window is well above half) all lanes step through
    // every event of the chunk, broadcast with readlane.
    const unsigned span = unsigned(2 * R);
    for (int base = 0; base < E; base += kWave) {
      const int cnt = min(kWave, E - base);
      const T vz = L < cnt ? sm.ez[base + L] : T(0);
      const int vp = L < cnt ? sm.ep[base + L] : 0;
      bool walk = false;
      uint64_t mine = 0;
      if (R <= kWalkMaxR) {
        mine = lane_event_mask(sm.colmask, vp >> 2, L < cnt, wn, R);
        walk = 4 * wave_max(popc(mine)) < 3 * cnt;
      }
      if (walk) {
        for (;;) {
          const bool on = mine != 0;
          if (ballot(on) == 0) break;
          const int i = on ? int(__builtin_ctzll(mine)) : 0;
          mine &= mine - 1;
          const int pk = __shfl(vp, i);
          const T z = __shfl(vz, i);
          st.apply(z, pk & 3, (pk >> 2) - L, on, active, g, stage);
        }
      } else {
        for (int i = 0; i < cnt; ++i) {
          const int pk = __builtin_amdgcn_readlane(vp, i);
          const int k = (pk >> 2) - L;
          st.apply(lane_bcast(vz, i), pk & 3, k, unsigned(k) <= span, active, g, stage);
        }
      }
    }
    // equal z never straddles a slab (slabs end at bin boundaries)
    if (ballot(st.pend != 0) != 0) st.close_points(st.pend != 0, active, g, stage);
    b_lo = b_hi;
  }
  if (cells) {  // the cell walk's lane state is the result
    st.count = cw.count;
    st.ok = cw.ok;
    st.last_hi = cw.last_hi;
  }
  if (ballot(active && st.ok == 0) != 0) raise_status(g.status, kStatusSlotOverflow);
  if (active) g.m_cnt[count_index(g.ny, x, y0 + L)] = st.count;
  if (active && st.count > 0) g.c_hi[count_index(g.ny, x, y0 + L)] = st.last_hi;
  stage.finish(st.count, st.last_hi, active);
}

// Waves per SIMD the z-sweep is compiled for: four (at most 128 registers;
// the register-staged blocks take about 130 unbounded), or fewer where the
// LDS of a CU's 160 KB allows fewer (one workgroup of one wave each).
constexpr int zsweep_waves(size_t lds_bytes) {
  return int(160 * 1024 / lds_bytes) / 4 < 1 ? 1 : (int(160 * 1024 / lds_bytes) / 4 > 4 ? 4 : int(160 * 1024 / lds_bytes) / 4);
}

template <class T, int ECAP, int NB, int WMAX, int NW>
__global__ __launch_bounds__(kWave, zsweep_waves(sizeof(Pass1ZSmem<T, ECAP, NB, WMAX>))) void pass1_zsweep_kernel(Pass1ZArgs<T> g) {
  __shared__ Pass1ZSmem<T, ECAP, NB, WMAX> sm;
  zsweep_tile<T, ECAP, NB, WMAX, NW>(g, xcd_tile(blockIdx.x, gridDim.x), sm);
}

// Batched: job j owns the global tiles [tile_base[j], tile_base[j + 1]);
// this launch covers the n tiles from t0 on (one radius class), one per
// workgroup, or (kQueue) as a persistent grid (take_tiles), as in
// pass2_gather_batch_kernel.
template <class T, int ECAP, int NB, int WMAX, int NW, bool kQueue>
__global__ __launch_bounds__(kWave, zsweep_waves(sizeof(Pass1ZSmem<T, ECAP, NB, WMAX>))) void pass1_zsweep_batch_kernel(const Pass1ZArgs<T>* jobs,
                                                                   const int64_t* tile_base, int njobs,
                                                                   int64_t t0, int64_t n, unsigned* queue) {
  __shared__ Pass1ZSmem<T, ECAP, NB, WMAX> sm;
  auto body = [&](int64_t i) {
    const int64_t t = t0 + i;
    const int j = find_job(tile_base, njobs, t);
    const Pass1ZArgs<T> g = jobs[j];
    zsweep_tile<T, ECAP, NB, WMAX, NW>(g, t - tile_base[j], sm);
  };
  if constexpr (kQueue)
    take_tiles(queue, n, body);
  else
    body(xcd_chunk_tile(blockIdx.x, n));
}

}  // namespace vo
