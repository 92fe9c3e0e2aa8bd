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
  T* m_lo;
  T* m_hi;
  uint16_t* m_dy;
  int64_t arena;     // triples reserved per chunk in the output arena
  int64_t* m_start;  // per column: first triple in the arena
  int32_t* m_cnt;    // per column: number of triples
  int* status;
};

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
  const int64_t obase = int64_t(blockIdx.x) * g.arena;
  // forward sweep, k = y
  int cur = 0, n = 0, apos = 0;
  bool ok = true;
  for (int y = max(0, y0 - g.rint); y < y1 && ok; ++y) {
    const int64_t col = int64_t(x) + int64_t(g.nx) * y;
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
  int64_t opos = 0;
  for (int y = min(g.ny - 1, y1 - 1 + g.rint); y >= y0; --y) {
    const int k = g.ny - 1 - y;
    const int64_t col = int64_t(x) + int64_t(g.nx) * y;
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
    if (nu < 0 || opos + nu > g.arena) {
      raise_status(g.status, nu < 0 ? kStatusP1Union : kStatusArenaOverflow);
      return;
    }
    for (int i = L; i < nu; i += kWave) {
      g.m_lo[obase + opos + i] = sm.ulo[i];
      g.m_hi[obase + opos + i] = sm.uhi[i];
      g.m_dy[obase + opos + i] = uint16_t(-sm.ukey[i]);
    }
    if (L == 0) {
      g.m_start[col] = obase + opos;
      g.m_cnt[col] = nu;
    }
    opos += nu;
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
  const T* m_lo;
  const T* m_hi;
  const uint16_t* m_dy;
  const int64_t* m_start;
  const int32_t* m_cnt;
  int64_t m_total;   // capacity of the pass-1 arena (bounds every read)
  T* o_z;            // output arena (endpoints)
  int64_t arena;     // endpoints reserved per chunk
  int64_t* o_start;  // per column
  int64_t* o_cnt;    // per column, in endpoints
  int* status;
  // list mode: recompute only the gather tiles listed in tiles[1 .. tiles[0]]
  // (kGatherTile columns each, split into chunks of `chunk` columns), with
  // arenas from o_base on
  const int* tiles;
  int64_t o_base;
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
__device__ int pass2_step(S& sm, int& cs, int& ns, int& cp, int& np, int k, int64_t mstart, int mcnt,
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
  if (mstart < 0 || mcnt < 0 || mstart + mcnt > g.m_total) {
    raise_status(g.status, kStatusArenaOverflow);
    return -1;
  }
  for (int b0 = 0; b0 < mcnt; b0 += ncap) {
    const int nb = min(ncap, mcnt - b0);
    for (int i = L; i < nb; i += kWave) {
      const int dy = g.m_dy[mstart + b0 + i];
      sm.nlo[i] = g.m_lo[mstart + b0 + i];
      sm.nhi[i] = g.m_hi[mstart + b0 + i];
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
  if (g.tiles) {
    const int per = kGatherTile / g.chunk;
    const int tile = g.tiles[1 + blockIdx.x / per];
    y = tile % g.ny;
    x0 = (tile / g.ny) * kGatherTile + (blockIdx.x % per) * g.chunk;
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
    const int64_t col = int64_t(x) + int64_t(g.nx) * y;
    const int nu = pass2_step<T>(sm, cs, ns, cp, np, x, g.m_start[col], g.m_cnt[col], g, NCAP, SCAP,
                                 PCAP, OCAP, sm.ulo[2], sm.uhi[2]);
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
    const int64_t col = int64_t(x) + int64_t(g.nx) * y;
    const int nu = pass2_step<T>(sm, cs, ns, cp, np, k, g.m_start[col], g.m_cnt[col], g, NCAP, SCAP,
                                 PCAP, OCAP, sm.ulo[2], sm.uhi[2]);
    if (nu < 0) return;
    if (x >= x1) continue;
    const int fa = sm.fpos[x - x0], fn = sm.fcnt[x - x0];
    const int nr = union_lists(sm.flo + fa, sm.fhi + fa, fn, sm.ulo[2], sm.uhi[2], nu, sm.mlo, sm.mhi,
                               2 * OCAP, sm.ulo[0], sm.uhi[0], OCAP);
    if (nr < 0 || opos + 2 * nr > g.arena) {
      raise_status(g.status, nr < 0 ? kStatusP2Slice : kStatusArenaOverflow);
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

// Gather per-column results (start, count) into CSR at precomputed offsets.
template <class T>
__global__ void gather_kernel(int64_t ncols, const T* src, const int64_t* start, const int64_t* cnt,
                              const int64_t* off, T* dst) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int64_t s = start[c], o = off[c];
  const int64_t n = cnt[c];
  for (int64_t i = 0; i < n; ++i) dst[o + i] = src[s + i];
}

// Complement on a grid grown by a ring of full columns
// (VoronoiMorpho::negate, src/vor3d/Voronoi.cpp:18-55, negate_ray
// MorphologyOperators.cpp:230-259).  Phase 0 writes lengths, phase 1 data.
template <class T>
__global__ void ring_complement_kernel(int nx, int ny, const int64_t* off, const T* z, T lo, T hi,
                                       int phase, int64_t* out_len, const int64_t* out_off, T* out) {
  const int64_t C = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int NX = nx + 2, NY = ny + 2;
  if (C >= int64_t(NX) * NY) return;
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

// Crop the ring and complement within [lo, hi] with parity counting
// (negateInv, Voronoi.cpp:57-89; negate_ray_range, MorphologyOperators.cpp:282-315).
template <class T>
__global__ void crop_complement_kernel(int nx, int ny, const int64_t* off, const T* z, T lo, T hi,
                                       int phase, int64_t* out_len, const int64_t* out_off, T* out) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= int64_t(nx) * ny) return;
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

}  // namespace vo

// ---------------------------------------------------------------------------
// Pass 2, gather form.  The output at column x is the union over source
// columns x' with |x - x'| = d <= R of the pass-1 pieces dilated by
// H = sqrt(R^2 - dy^2 - d^2): exactly what SeparatePower2D's forward and
// backward sweeps assemble (their getLine emits [y1 - dz, y2 + dz] for every
// live interior and point, SeparatePower2D.cpp:295-341).  One lane per output
// column; the pieces of the tile's source window are staged in LDS once.
// Within one source column the pieces are z-sorted, so touching dilated pieces
// fuse into runs that are united into the lane's sorted output list with a
// forward-moving cursor.  A tile whose window or output lists do not fit is
// flagged and recomputed by the sweep kernel.
// ---------------------------------------------------------------------------

namespace vo {

template <class T>
struct Pass2GArgs {
  int nx, ny;
  int rint;
  double R2;
  int ntx;  // tiles per row
  const T* m_lo;
  const T* m_hi;
  const uint16_t* m_dy;
  const int64_t* m_start;
  const int32_t* m_cnt;
  int64_t m_total;
  T* o_z;
  int64_t arena;  // endpoints reserved per tile
  int64_t* o_start;
  int64_t* o_cnt;
  int* tile_list;  // [0] = count, then the ids of tiles to recompute with the sweep kernel
  int ntiles;
  int* status;
};


template <class T, int POOL, int OCAP, int WMAX>
struct Pass2GSmem {
  T pa[POOL];
  T pb[POOL];
  uint8_t pdy[POOL];
  int coff[WMAX + 1];
  double hrow[256];
  double olo[OCAP][kWave];
  double ohi[OCAP][kWave];
};

// Unites [s, e] into the lane's sorted disjoint list (touching fuses).  `k`
// is a cursor that only needs to be at or before the insertion point for a
// sequence of increasing s; it is moved back when needed.  Returns false on
// overflow.
template <int OCAP>
__device__ __forceinline__ bool lane_unite(double (*olo)[kWave], double (*ohi)[kWave], int L, int& n, int& k,
                                           double s, double e) {
  while (k > 0 && ohi[k - 1][L] >= s) --k;
  while (k < n && ohi[k][L] < s) ++k;
  if (k == n || olo[k][L] > e) {
    if (n == OCAP) return false;
    for (int j = n; j > k; --j) {
      olo[j][L] = olo[j - 1][L];
      ohi[j][L] = ohi[j - 1][L];
    }
    olo[k][L] = s;
    ohi[k][L] = e;
    ++n;
    return true;
  }
  if (s < olo[k][L]) olo[k][L] = s;
  double hi = fmax(ohi[k][L], e);
  int j = k + 1;
  while (j < n && olo[j][L] <= hi) {
    hi = fmax(hi, ohi[j][L]);
    ++j;
  }
  const int gone = j - k - 1;
  if (gone > 0) {
    for (int t = k + 1; t + gone < n; ++t) {
      olo[t][L] = olo[t + gone][L];
      ohi[t][L] = ohi[t + gone][L];
    }
    n -= gone;
  }
  ohi[k][L] = hi;
  return true;
}

template <class T, int POOL, int OCAP, int WMAX>
__global__ __launch_bounds__(kWave) void pass2_gather_kernel(Pass2GArgs<T> g) {
  using S = Pass2GSmem<T, POOL, OCAP, WMAX>;
  __shared__ S sm;
  const int L = lane_id();
  const int y = blockIdx.x % g.ny;
  const int t = blockIdx.x / g.ny;
  const int x0 = t * kGatherTile;
  if (x0 >= g.nx) return;
  if (*g.status != 0) return;  // pass 1 did not complete
  const int rint = g.rint;
  const int wlo = x0 - rint, wn = kGatherTile + 2 * rint;
  const int64_t rowbase = int64_t(g.nx) * y;
  // window column sizes -> prefix offsets into the pool
  int carry = 0;
  for (int base = 0; base < wn; base += kWave) {
    const int w = base + L;
    const int c = wlo + w;
    int cnt = 0;
    if (w < wn && c >= 0 && c < g.nx) cnt = g.m_cnt[rowbase + c];
    int v = cnt;  // inclusive scan of counts
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int o = __shfl_up(v, d);
      if (L >= d) v += o;
    }
    if (w < wn) sm.coff[w + 1] = carry + v;
    carry += __shfl(v, kWave - 1);
  }
  if (L == 0) sm.coff[0] = 0;
  wave_sync();
  const int total = sm.coff[wn];
  if (total > POOL) {
    if (L == 0) g.tile_list[1 + atomicAdd(g.tile_list, 1)] = blockIdx.x;
    return;
  }
  // stage the window's pieces: one lane per source column
  for (int w = L; w < wn; w += kWave) {
    const int c = wlo + w;
    if (c < 0 || c >= g.nx) continue;
    const int64_t st = g.m_start[rowbase + c];
    const int o = sm.coff[w], m = sm.coff[w + 1] - o;
    if (st < 0 || st + m > g.m_total) {
      raise_status(g.status, kStatusArenaOverflow);
      continue;
    }
    for (int i = 0; i < m; ++i) {
      sm.pa[o + i] = g.m_lo[st + i];
      sm.pb[o + i] = g.m_hi[st + i];
      sm.pdy[o + i] = uint8_t(g.m_dy[st + i]);
    }
  }
  wave_sync();
  const int x = x0 + L;
  const bool active = x < g.nx;
  int n = 0;
  bool ok = true;
  // source columns nearest first: d = 0, -1, +1, -2, +2, ...
  for (int step = 0; step <= 2 * rint; ++step) {
    const int ad = (step + 1) >> 1;
    const int d = (step & 1) ? -ad : ad;
    if (step == 0 || (step & 1)) {  // new |d|: table of H by dy
      wave_sync();
      for (int dy = L; dy <= rint; dy += kWave) {
        const double v = g.R2 - double(dy) * double(dy) - double(ad) * double(ad);
        sm.hrow[dy] = v >= 0 ? sqrt(v) : -1.0;
      }
      wave_sync();
    }
    const int c = x + d;
    if (!active || !ok || c < 0 || c >= g.nx) continue;
    const int w = c - wlo;
    const int p0 = sm.coff[w], p1 = sm.coff[w + 1];
    int k = 0;
    bool run = false;
    double rs = 0, re = 0;
    for (int p = p0; p < p1; ++p) {
      const double h = sm.hrow[sm.pdy[p]];
      if (h < 0) continue;
      const double s = double(sm.pa[p]) - h, e = double(sm.pb[p]) + h;
      if (run && s <= re) {
        rs = fmin(rs, s);
        re = fmax(re, e);
        continue;
      }
      if (run && !lane_unite<OCAP>(sm.olo, sm.ohi, L, n, k, rs, re)) {
        ok = false;
        break;
      }
      run = true;
      rs = s;
      re = e;
    }
    if (ok && run && !lane_unite<OCAP>(sm.olo, sm.ohi, L, n, k, rs, re)) ok = false;
  }
  if (ballot(active && !ok) != 0) {
    if (L == 0) g.tile_list[1 + atomicAdd(g.tile_list, 1)] = blockIdx.x;
    return;
  }
  if (!active) return;
  const int64_t col = rowbase + x;
  const int64_t ob = int64_t(blockIdx.x) * g.arena + int64_t(L) * 2 * OCAP;
  for (int i = 0; i < n; ++i) {
    g.o_z[ob + 2 * i] = T(sm.olo[i][L]);
    g.o_z[ob + 2 * i + 1] = T(sm.ohi[i][L]);
  }
  g.o_start[col] = ob;
  g.o_cnt[col] = 2 * n;
}

}  // namespace vo
