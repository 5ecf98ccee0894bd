// deblock.hip — the encoder's loop filter on gfx950.
//
// Filters a reconstructed picture in place as the encoder does before the picture becomes a
// reference (enc/mod.rs:653-654; h264/deblock.rs:232-421 and :608-695): frame macroblocks, 4:2:0,
// 8-bit samples, every slice filtered across slice edges with zero alpha/beta and chroma QP offsets
// (the encoder's slice parameters), no I_PCM or lossless macroblocks.
//
// Macroblocks are filtered in raster order, each its vertical edges and then its horizontal ones,
// and a macroblock's edges read what the macroblocks left of and above it wrote. One wave filters
// one macroblock row from left to right. Macroblock (x, y) starts once row y-1 has finished
// macroblock x+1, whose left edge changes the right columns of (x, y-1): the wavefront the
// reference's `deblock_picture_par` runs on threads (h264/deblock.rs:583-603). Rows publish their
// progress with an agent-scope release store on a counter per row and read it with an acquire load
// (rows of one picture run on different CUs and XCDs); a row re-reads its counter only when the last
// value it saw is not far enough ahead.
//
// Within a macroblock, lanes 0-15 hold the 16 luma lines, lanes 16-23 the Cb lines and 24-31 the Cr
// lines. For the vertical edges a lane holds its row (the macroblock's samples and the 4 columns left
// of it) in registers; an LDS tile transposes the rows into columns (with the 4 rows above) for the
// horizontal edges, and back for the stores. Lanes 0-31 first derive the 32 boundary strengths
// (direction x edge x 4-line segment) of the macroblock, one each, and the filtering lanes gather
// theirs with cross-lane reads.
#include <mutex>

#include <cstddef>
#include "common.h"
#include "h264_tables.h"

namespace ia {
namespace {

constexpr int SPIN_LIMIT = 1 << 20;  // polls of the row above before the kernel reports a stall

__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }
__device__ __forceinline__ int clip3(int lo, int hi, int v) { return min(max(v, lo), hi); }

// Whether the transform block holding 4x4 block (bx, by) has coefficients (h264/deblock.rs:336-344).
__device__ __forceinline__ bool coded(const ia_mbinfo &m, int bx, int by) {
  if (m.t8x8) {
    const int x0 = bx & 2, y0 = by & 2;
    return (m.nnz[y0 * 4 + x0] | m.nnz[y0 * 4 + x0 + 1] | m.nnz[(y0 + 1) * 4 + x0] | m.nnz[(y0 + 1) * 4 + x0 + 1]) != 0;
  }
  return m.nnz[by * 4 + bx] != 0;
}

// Vectors a whole sample or more apart (frame macroblocks: the vertical limit is 4 quarter samples).
__device__ __forceinline__ bool far(int ax, int ay, int bx, int by) { return iabs(ax - bx) >= 4 || iabs(ay - by) >= 4; }

// Different references, a different number of vectors, or vectors a whole sample apart
// (h264/deblock.rs:347-383). List 1 is unused in P pictures: reference -1, vector 0.
__device__ __forceinline__ bool motion_differs(const ia_mbinfo &p, int pbx, int pby, const ia_mbinfo &q, int qbx,
                                               int qby) {
  const int p8 = (pby >> 1) * 2 + (pbx >> 1), q8 = (qby >> 1) * 2 + (qbx >> 1);
  const int pr0 = p.ref_id[p8], qr0 = q.ref_id[q8];
  const int px = p.mv[pby * 4 + pbx][0], py = p.mv[pby * 4 + pbx][1];
  const int qx = q.mv[qby * 4 + qbx][0], qy = q.mv[qby * 4 + qbx][1];
  const int np = pr0 >= 0, nq = qr0 >= 0;
  if (np != nq) return true;
  if (np == 1) return pr0 != qr0 || far(px, py, qx, qy);
  // no vector on either side (list 1's vectors are 0): the reference's last case
  return far(px, py, qx, qy) && (far(px, py, 0, 0) || far(0, 0, qx, qy));
}

// Boundary strength of segment i of edge e (0: the macroblock edge) in direction dir (0: vertical
// edges) of macroblock m; o is the macroblock across a macroblock edge (h264/deblock.rs:391-408).
__device__ __forceinline__ int strength(const ia_mbinfo &m, const ia_mbinfo &o, int dir, int e, int i) {
  if (m.kind != IA_KIND_INTER || o.kind != IA_KIND_INTER) return e == 0 ? 4 : 3;
  const int qbx = dir == 0 ? e : i, qby = dir == 0 ? i : e;
  const int pbx = dir == 0 ? (e == 0 ? 3 : e - 1) : i, pby = dir == 0 ? i : (e == 0 ? 3 : e - 1);
  if (coded(m, qbx, qby) || coded(o, pbx, pby)) return 2;
  return motion_differs(o, pbx, pby, m, qbx, qby) ? 1 : 0;
}

struct Thr {
  int alpha, beta, tc0;
};

// Orders the LDS and memory accesses of the lanes of a one-wave workgroup: the wave executes its
// LDS accesses in order, so only the compiler must not move accesses across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One sample line across an edge, p3..p0 | q0..q3, strength b (0: no filtering), without branches
// so that luma and chroma lanes share the instructions (h264/deblock.rs:279-334). Chroma lines use
// p1..q1 only: their p2/q2 updates are off (as if |p2 - p0| and |q2 - q0| were at least beta), tc is
// tc0 + 1, and strength 4 takes the 3-tap p0/q0 formulas, which are luma's outside its strong case.
__device__ __forceinline__ void filter_line(int &p3, int &p2, int &p1, int &p0, int &q0, int &q1, int &q2, int &q3,
                                            int b, int alpha, int beta, int tc0, bool chroma) {
  const int P0 = p0, Q0 = q0, P1 = p1, Q1 = q1, P2 = p2, Q2 = q2;
  const bool on = b > 0 && iabs(P0 - Q0) < alpha && iabs(P1 - P0) < beta && iabs(Q1 - Q0) < beta;
  const bool ap = !chroma && iabs(P2 - P0) < beta, aq = !chroma && iabs(Q2 - Q0) < beta;
  const bool strong = b == 4;
  int n0 = P0, m0 = Q0, n1 = P1, m1 = Q1, n2 = P2, m2 = Q2;
  // each form only where some lane of the wave takes it (a wave-uniform branch)
  if (__any(on && !strong)) {  // strength 1-3
    const int tc = chroma ? tc0 + 1 : tc0 + ap + aq;
    const int delta = clip3(-tc, tc, ((Q0 - P0) * 4 + (P1 - Q1) + 4) >> 3);
    const int avg = (P0 + Q0 + 1) >> 1;
    if (!strong) {
      n0 = clip3(0, 255, P0 + delta);
      m0 = clip3(0, 255, Q0 - delta);
      n1 = ap ? P1 + clip3(-tc0, tc0, (P2 + avg - 2 * P1) >> 1) : P1;
      m1 = aq ? Q1 + clip3(-tc0, tc0, (Q2 + avg - 2 * Q1) >> 1) : Q1;
    }
  }
  if (__any(on && strong)) {  // strength 4
    const bool small = iabs(P0 - Q0) < (alpha >> 2) + 2;
    const bool sp = ap && small, sq = aq && small;
    if (strong) {
      n0 = sp ? (P2 + 2 * P1 + 2 * P0 + 2 * Q0 + Q1 + 4) >> 3 : (2 * P1 + P0 + Q1 + 2) >> 2;
      n1 = sp ? (P2 + P1 + P0 + Q0 + 2) >> 2 : P1;
      n2 = sp ? (2 * p3 + 3 * P2 + P1 + P0 + Q0 + 4) >> 3 : P2;
      m0 = sq ? (P1 + 2 * P0 + 2 * Q0 + 2 * Q1 + Q2 + 4) >> 3 : (2 * Q1 + Q0 + P1 + 2) >> 2;
      m1 = sq ? (P0 + Q0 + Q1 + Q2 + 2) >> 2 : Q1;
      m2 = sq ? (2 * q3 + 3 * Q2 + Q1 + Q0 + P0 + 4) >> 3 : Q2;
    }
  }
  if (on) {
    p0 = n0, q0 = m0, p1 = n1, q1 = m1, p2 = n2, q2 = m2;
  }
}

__device__ __forceinline__ uint32_t pack4(const int *v) {
  return (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
}
__device__ __forceinline__ void unpack4(uint32_t w, int *v) {
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = (w >> (8 * k)) & 0xff;
}

// Tiles: luma rows y0-4 .. y0+15 with columns x0-4 .. x0+15 at 12 .. 31 (the macroblock's own 16
// columns 16-aligned); chroma rows cy0-2 .. cy0+7 with columns cx0-4 .. cx0+7 at 4 .. 15.
constexpr int LX = 12, CX = 4;
struct DbLds {
  __attribute__((aligned(16))) uint8_t y[20][32];
  __attribute__((aligned(16))) uint8_t c[2][10][16];
  uint8_t bs[32];     // strengths [dir][e][segment]
  ia_mbinfo rm[2];    // records of the row's macroblocks x (at x & 1) and x - 1
  ia_mbinfo rt[2];    // record of the macroblock above x (at x & 1)
  uint8_t alpha[52], beta[52], qpc[52];  // the filter's tables (and chroma QP per QP), copied once
  int8_t tc0[52][4];
};
// alpha, beta and tc0 of strength b at average QP qpav (offsets 0: indexA = indexB = qpav), from the
// LDS copies of the tables.
__device__ __forceinline__ Thr thresholds(const DbLds &L, int qpav, int b) {
  const int ia = clip3(0, 51, qpav);
  Thr t;
  t.alpha = L.alpha[ia];
  t.beta = L.beta[ia];
  t.tc0 = max((int)L.tc0[ia][b & 3], 0);  // read for every strength; used for 1-3 only
  if (t.beta == 0) t.alpha = 0;           // no filtering: alpha 0 makes |p0 - q0| < alpha false
  return t;
}

__device__ __forceinline__ int chroma_qp(const DbLds &L, int qp) { return L.qpc[clip3(0, 51, qp)]; }

// The filter runs on P and I pictures (the encoder's references; B pictures are never kept): it reads
// the records up to the list 0 motion, their first REC_DWORDS dwords, and leaves list 1 as in P pictures.
constexpr int REC_STRIDE = sizeof(ia_mbinfo) / 4;
constexpr int REC_DWORDS = offsetof(ia_mbinfo, mv1) / 4;
static_assert(sizeof(ia_mbinfo) % 4 == 0 && offsetof(ia_mbinfo, mv1) % 4 == 0 && REC_DWORDS <= 32,
              "a record's list 0 part is read as dwords by lanes 0-31 / 32-63");

// Macroblock row `row` of one picture (one wave). prog[r]: macroblocks of row r finished (the
// counters start at 0 and are 0 again when the launch ends); err: set when a row waited too long.
__device__ __forceinline__ void deblock_row(DbLds &L, uint8_t *__restrict__ py, uint8_t *__restrict__ pcb, uint8_t *__restrict__ pcr,
                            const ia_mbinfo *__restrict__ mbs, int mb_w, int mb_h, int row, int *prog, int *err) {
  const int lane = threadIdx.x;
  const int ls = mb_w * 16, cs = mb_w * 8;
  const int y0 = row * 16, cy0 = row * 8;
  const bool luma = lane < 16, chroma = lane >= 16 && lane < 32;
  const int cl = (lane - 16) & 7;            // chroma line
  uint8_t *cpl = lane < 24 ? pcb : pcr;      // chroma plane of lanes 16-31
  const int comp = lane < 24 ? 0 : 1;
  int left[4] = {0, 0, 0, 0};                // the 4 samples left of the macroblock on the lane's line
  int seen = 0;                              // last progress of the row above this wave read
  // A macroblock's records and own samples are not changed by the filter: the wave fetches those of
  // macroblock x+1 while it filters x (records into LDS, samples into registers).
  const uint32_t *grec = reinterpret_cast<const uint32_t *>(mbs);
  // a lane's dword of the records of macroblock x (lanes 0-31) or of the one above it (32-63)
  auto fetch_record = [&](int x) -> uint32_t {
    const int k = lane & 31;
    if (k >= REC_DWORDS || (lane >= 32 && row == 0)) return 0u;
    return grec[(size_t)((lane < 32 ? row : row - 1) * mb_w + x) * REC_STRIDE + k];
  };
  auto store_record = [&](int x, uint32_t v) {
    const int k = lane & 31;
    if (k < REC_DWORDS) reinterpret_cast<uint32_t *>(lane < 32 ? &L.rm[x & 1] : &L.rt[x & 1])[k] = v;
  };
  auto fetch_samples = [&](int x, uint4 &v) {
    if (luma) v = *reinterpret_cast<const uint4 *>(py + (size_t)(y0 + lane) * ls + x * 16);
    else if (chroma) {
      const uint2 c = *reinterpret_cast<const uint2 *>(cpl + (size_t)(cy0 + cl) * cs + x * 8);
      v.x = c.x;
      v.y = c.y;
    }
  };
  for (int q = lane; q < 52; q += 64) {
    L.alpha[q] = IA_ALPHA[q];
    L.beta[q] = IA_BETA[q];
    L.qpc[q] = (uint8_t)(q < 30 ? q : IA_CHROMA_QP[q - 30]);
#pragma unroll
    for (int k = 0; k < 4; k++) L.tc0[q][k] = IA_TC0[q][k];
  }
  uint4 nv = make_uint4(0, 0, 0, 0);
  uint32_t nrec = fetch_record(0);
  fetch_samples(0, nv);
  store_record(0, nrec);
  __syncthreads();
  for (int x = 0; x < mb_w; x++) {
    const int x0 = x * 16, cx0 = x * 8;
    // the macroblock's own samples, fetched during the previous macroblock
    int s[20];  // luma lane: columns x0-4 .. x0+15 of row y0+lane; chroma lane: columns cx0-4 .. cx0+7
    // luma lanes: s[4 + k] = column x0 + k. Chroma lanes: columns cx0 .. cx0+7 at s[4, 5, 10 .. 15],
    // so that the chroma edges at cx0 and cx0+4 sit where the luma edges at x0 and x0+8 do
    if (chroma) {
      int t[8];
      unpack4(nv.x, t);
      unpack4(nv.y, t + 4);
      s[4] = t[0], s[5] = t[1], s[6] = s[7] = s[8] = s[9] = 0;
#pragma unroll
      for (int k = 2; k < 8; k++) s[8 + k] = t[k];
      s[16] = s[17] = s[18] = s[19] = 0;
    } else {
      unpack4(nv.x, s + 4);
      unpack4(nv.y, s + 8);
      unpack4(nv.z, s + 12);
      unpack4(nv.w, s + 16);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) s[k] = left[k];
    const ia_mbinfo &m = L.rm[x & 1];
    const ia_mbinfo &ml = x > 0 ? L.rm[(x - 1) & 1] : m;
    const ia_mbinfo &mt = row > 0 ? L.rt[x & 1] : m;
    const bool live = m.slice != 0;
    const bool left_ok = live && x > 0 && ml.slice != 0;
    const bool top_ok = live && row > 0 && mt.slice != 0;
    // boundary strengths, one per lane 0-31: dir = lane >> 4, edge e = (lane >> 2) & 3, segment lane & 3
    // an inter macroblock that moves as a whole and has no coefficients leaves its inner edges
    // unfiltered (h264/deblock.rs:628-634, :385-389): lanes 0-15 check a 4x4 block each, 16-19 an 8x8 block
    bool same = true;
    if (lane < 16) same = m.nnz[lane] == 0 && m.mv[lane][0] == m.mv[0][0] && m.mv[lane][1] == m.mv[0][1];
    else if (lane < 20) same = m.ref_idx[lane - 16] == m.ref_idx[0] && m.ref_id[lane - 16] == m.ref_id[0];
    const bool still = m.kind == IA_KIND_INTER && __all(same);
    if (lane < 32 && live) {
      const int dir = lane >> 4, e = (lane >> 2) & 3, i = lane & 3;
      L.bs[lane] = (uint8_t)(e > 0 && still ? 0 : strength(m, e == 0 ? (dir == 0 ? ml : mt) : m, dir, e, i));
    }
    // wait for macroblock x+1 of the row above (the last one at the row's end); after a wait that
    // timed out the row waits no more (the launch reports the stall and its output is not used)
    if (row > 0) {
      const int need = min(x + 2, mb_w);
      if (seen < need) {
        if (lane == 0) {
          int v = __hip_atomic_load(prog + row - 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          for (int n = 0; v < need; n++) {
            if (n == SPIN_LIMIT) {
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              v = mb_w;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            v = __hip_atomic_load(prog + row - 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          }
          seen = v;
        }
        seen = __shfl(seen, 0);
      }
    }
    wave_sync();  // strengths in LDS; the row above's stores visible to every lane (the acquire is the wave's)
    // the rows above: luma y0-4 .. y0-1 into tile rows 0-3, chroma cy0-2 .. cy0-1 into rows 0-1
    if (top_ok) {
      if (lane < 4) {
        const uint4 v = *reinterpret_cast<const uint4 *>(py + (size_t)(y0 - 4 + lane) * ls + x0);
        *reinterpret_cast<uint4 *>(&L.y[lane][LX + 4]) = v;
      } else if (lane >= 16 && lane < 20) {
        const int c = (lane - 16) >> 1, r = (lane - 16) & 1;
        const uint2 v = *reinterpret_cast<const uint2 *>((c ? pcr : pcb) + (size_t)(cy0 - 2 + r) * cs + cx0);
        *reinterpret_cast<uint2 *>(&L.c[c][r][CX + 4]) = v;
      }
    }
    // fetch macroblock x+1's records and samples now: they arrive while x is filtered (the release
    // at the end of x waits for every load of the wave)
    if (x + 1 < mb_w) {
      nrec = fetch_record(x + 1);
      fetch_samples(x + 1, nv);
    }
    // per lane: the strength segment of its line, and the QPs of its macroblock edge and inner edges
    const int seg = luma ? lane >> 2 : cl >> 1;
    const int qpm = luma ? m.qp : chroma_qp(L, m.qp);
    const int qpv = luma ? (ml.qp + m.qp + 1) >> 1 : (chroma_qp(L, ml.qp) + chroma_qp(L, m.qp) + 1) >> 1;
    const int qph = luma ? (mt.qp + m.qp + 1) >> 1 : (chroma_qp(L, mt.qp) + chroma_qp(L, m.qp) + 1) >> 1;
    const bool line = (luma || chroma) && live;
    // the strength of edge e for the lane (0: not filtered): luma skips the odd inner edges of an
    // 8x8-transform macroblock; chroma has edges 0 and 2 only (its sample 0 and 4)
    auto edge_bs = [&](int dir, int e, bool ok0) -> int {
      const bool use = line && (e > 0 || ok0) && (luma ? !(m.t8x8 && (e & 1)) : (e & 1) == 0);
      return use ? L.bs[dir * 16 + e * 4 + seg] : 0;
    };
    // vertical edges, on the lane's row in registers
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int b = edge_bs(0, e, left_ok);
      const Thr t = thresholds(L, e == 0 ? qpv : qpm, b);
      const int q = 4 + 4 * e;
      if (__any(b != 0))
        filter_line(s[q - 4], s[q - 3], s[q - 2], s[q - 1], s[q], s[q + 1], s[q + 2], s[q + 3], b, t.alpha, t.beta,
                    t.tc0, chroma);
    }
    // rows into the tiles (chroma lanes gather their spread samples back)
    if (luma) {
#pragma unroll
      for (int k = 0; k < 5; k++) *reinterpret_cast<uint32_t *>(&L.y[4 + lane][LX + 4 * k]) = pack4(s + 4 * k);
    } else if (chroma) {
      const int u[4] = {s[4], s[5], s[10], s[11]};
      *reinterpret_cast<uint32_t *>(&L.c[comp][2 + cl][CX]) = pack4(s);
      *reinterpret_cast<uint32_t *>(&L.c[comp][2 + cl][CX + 4]) = pack4(u);
      *reinterpret_cast<uint32_t *>(&L.c[comp][2 + cl][CX + 8]) = pack4(s + 12);
    }
    wave_sync();
    // horizontal edges, on the lane's column: luma rows y0-4 .. y0+15 at v[0 .. 19]; chroma rows
    // cy0-2 .. cy0+7 at v[2 .. 5, 10 .. 15] (its edges at v[4] and v[12], as luma's edges 0 and 2)
    int v[20];
    if (luma) {
#pragma unroll
      for (int r = 0; r < 20; r++) v[r] = L.y[r][LX + 4 + lane];
    } else {
      const uint8_t(*ct)[16] = L.c[comp];
#pragma unroll
      for (int r = 0; r < 20; r++) v[r] = 0;
      if (chroma) {
        v[2] = ct[0][CX + 4 + cl], v[3] = ct[1][CX + 4 + cl], v[4] = ct[2][CX + 4 + cl], v[5] = ct[3][CX + 4 + cl];
#pragma unroll
        for (int r = 4; r < 10; r++) v[6 + r] = ct[r][CX + 4 + cl];
      }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int b = edge_bs(1, e, top_ok);
      const Thr t = thresholds(L, e == 0 ? qph : qpm, b);
      const int q = 4 + 4 * e;
      if (__any(b != 0))
        filter_line(v[q - 4], v[q - 3], v[q - 2], v[q - 1], v[q], v[q + 1], v[q + 2], v[q + 3], b, t.alpha, t.beta,
                    t.tc0, chroma);
    }
    if (luma) {
#pragma unroll
      for (int r = 1; r < 20; r++) L.y[r][LX + 4 + lane] = (uint8_t)v[r];
    } else if (chroma) {
      uint8_t(*ct)[16] = L.c[comp];
      ct[1][CX + 4 + cl] = (uint8_t)v[3], ct[2][CX + 4 + cl] = (uint8_t)v[4], ct[3][CX + 4 + cl] = (uint8_t)v[5];
#pragma unroll
      for (int r = 4; r < 10; r++) ct[r][CX + 4 + cl] = (uint8_t)v[6 + r];
    }
    wave_sync();
    // stores: the lane's row (its 4 left samples belong to the macroblock on the left), the rows above
    if (luma) {
      uint32_t w[5];
#pragma unroll
      for (int k = 0; k < 5; k++) w[k] = *reinterpret_cast<const uint32_t *>(&L.y[4 + lane][LX + 4 * k]);
      uint8_t *rp = py + (size_t)(y0 + lane) * ls + x0;
      *reinterpret_cast<uint4 *>(rp) = make_uint4(w[1], w[2], w[3], w[4]);
      if (x > 0) *reinterpret_cast<uint32_t *>(rp - 4) = w[0];
      unpack4(w[4], left);
      if (top_ok && lane >= 1 && lane < 4)
        *reinterpret_cast<uint4 *>(py + (size_t)(y0 - 4 + lane) * ls + x0) = *reinterpret_cast<const uint4 *>(&L.y[lane][LX + 4]);
    } else if (chroma) {
      uint32_t w[3];
#pragma unroll
      for (int k = 0; k < 3; k++) w[k] = *reinterpret_cast<const uint32_t *>(&L.c[comp][2 + cl][CX + 4 * k]);
      uint8_t *rp = cpl + (size_t)(cy0 + cl) * cs + cx0;
      *reinterpret_cast<uint2 *>(rp) = make_uint2(w[1], w[2]);
      if (x > 0) *reinterpret_cast<uint32_t *>(rp - 4) = w[0];
      unpack4(w[2], left);
      if (top_ok && cl == 0)
        *reinterpret_cast<uint2 *>(cpl + (size_t)(cy0 - 1) * cs + cx0) = *reinterpret_cast<const uint2 *>(&L.c[comp][1][CX + 4]);
    }
    wave_sync();  // every lane's stores before the progress is published (the release waits for them)
    if (lane == 0) __hip_atomic_store(prog + row, x + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    // macroblock x+1's records replace x-1's (read no more)
    if (x + 1 < mb_w) store_record(x + 1, nrec);
    wave_sync();
  }
  // the counter of the row above is read no more: clear it for the next launch (and the last row its own)
  if (lane == 0) {
    if (row > 0) __hip_atomic_store(prog + row - 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (row == mb_h - 1) __hip_atomic_store(prog + row, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ __launch_bounds__(64) void deblock_kernel(uint8_t *py, uint8_t *pcb, uint8_t *pcr, const ia_mbinfo *mbs,
                                                      int mb_w, int mb_h, int *prog, int *err) {
  __shared__ DbLds lds;
  deblock_row(lds, py, pcb, pcr, mbs, mb_w, mb_h, blockIdx.x, prog, err);
}

// Many pictures: block b filters row b - first of the job whose rows hold it (jobs by `first`).
__global__ __launch_bounds__(64) void deblock_batch_kernel(const DbJob *jobs, int njobs, int *err) {
  __shared__ DbLds lds;
  const int b = blockIdx.x;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].first <= b) lo = mid;
    else hi = mid - 1;
  }
  const DbJob &j = jobs[lo];
  if (!j.y) return;
  deblock_row(lds, j.y, j.cb, j.cr, j.mbs, j.mb_w, j.mb_h, b - j.first, j.prog, err);
}

}  // namespace

hipError_t launch_deblock(uint8_t *y, uint8_t *cb, uint8_t *cr, const ia_mbinfo *mbs, int mb_w, int mb_h, int *prog,
                          int *err, hipStream_t stream) {
  if (mb_w < 1 || mb_h < 1 || !y || !cb || !cr || !mbs || !prog || !err) return hipErrorInvalidValue;
  hipLaunchKernelGGL(deblock_kernel, dim3(mb_h), dim3(64), 0, stream, y, cb, cr, mbs, mb_w, mb_h, prog, err);
  return hipGetLastError();
}

hipError_t launch_deblock_batch(const DbJob *jobs, int njobs, int nblocks, int *err, hipStream_t stream) {
  if (njobs < 1 || nblocks < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(deblock_batch_kernel, dim3(nblocks), dim3(64), 0, stream, jobs, njobs, err);
  return hipGetLastError();
}

}  // namespace ia

// ---- the loop filter without allocation ------------------------------------------------------------
//
// For callers that cannot allocate device memory (code under graph capture, tight serving loops): the
// row counters and the stall flag live in a file-scope device array instead of a context's buffers.
// Every call on a GPU shares that GPU's array, so calls that overlap on one GPU race.

namespace ia {
namespace {

constexpr int NOALLOC_MAX_ROWS = 512;  // pictures up to 8192 samples high

// [0] the stall flag, [1 + r] the progress counter of macroblock row r
__device__ __align__(256) int scratch[1 + NOALLOC_MAX_ROWS];

// the array's device address per device, from hipGetSymbolAddress on the first call there
constexpr int NOALLOC_MAX_DEVICES = 64;
int *g_scratch[NOALLOC_MAX_DEVICES];
std::mutex g_scratch_mu;

__global__ void zero_words(int *p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void copy_word(const int *src, int *dst) {
  if (threadIdx.x == 0) *dst = *src;
}

}  // namespace

}  // namespace ia

extern "C" int ia_deblock_picture_noalloc(uint8_t *y, uint8_t *cb, uint8_t *cr, const ia_mbinfo *mbs, int mb_w, int mb_h,
                                          int *stalled, void *stream) {
  if (!y || !cb || !cr || !mbs || mb_w < 1 || mb_h < 1 || mb_w > 512) return IA_EINVAL;
  if (mb_h > ia::NOALLOC_MAX_ROWS) return IA_EINVAL;  // beyond the static scratch array
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= ia::NOALLOC_MAX_DEVICES) return IA_EDEVICE;
  int *p;
  {
    std::lock_guard<std::mutex> lock(ia::g_scratch_mu);
    if (!ia::g_scratch[dev]) {
      void *addr = nullptr;
      if (hipGetSymbolAddress(&addr, HIP_SYMBOL(ia::scratch)) != hipSuccess || !addr) return IA_EDEVICE;
      ia::g_scratch[dev] = static_cast<int *>(addr);
    }
    p = ia::g_scratch[dev];
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  // the counters and the flag start at zero (a fill kernel, so that a captured graph replays it)
  hipLaunchKernelGGL(ia::zero_words, dim3(1), dim3(256), 0, st, p, 1 + mb_h);
  if (hipGetLastError() != hipSuccess) return IA_EDEVICE;
  if (ia::launch_deblock(y, cb, cr, mbs, mb_w, mb_h, p + 1, p, st) != hipSuccess) return IA_EDEVICE;
  if (stalled) {
    hipLaunchKernelGGL(ia::copy_word, dim3(1), dim3(64), 0, st, p, stalled);
    if (hipGetLastError() != hipSuccess) return IA_EDEVICE;
  }
  return IA_OK;
}
