// Device side of the graph-to-sequence alignment: the DP over a
// rank-ordered POA graph and the on-device traceback, for CDNA4 (gfx950).
//
// Semantics follow the reference SIMD engine
// (src/simd_alignment_engine_implementation.hpp of spoa v4.0.8, AVX2 lane
// counts): rows of W = ceil(n / L) * L columns with L = 16 (int16 scores)
// or L = 8 (int32 scores), padding columns scored with
// -max(|m|, |n|, |g|, |q|) (:526-542), linear / affine / convex
// recurrences (:824-901, :1179-1272, :1638-1774), the end-cell choice
// (:903-953) and the traceback with its tie order (:955-1124, :1324-1574,
// :1826-2121).  In convex mode a gap in E/Q opens from the pre-gap score
// H1 inside an L-column segment and from the final score at a segment
// start (:1738-1768); the kernel keeps that segment rule exactly.
//
// Work decomposition.  One workgroup aligns one (graph, read) job at a
// time and takes jobs from a queue (a persistent grid).  Lanes own K
// consecutive columns (K a multiple of L in convex mode, so every lane
// starts a segment); a row is one tile of nwaves * 64 * K columns or a
// few of them.  Each row runs
//   1. the vertical pass: every predecessor row (H, F, O) is read back
//      with 16-byte loads; the diagonal H_p[c0 - 1] comes from the
//      neighbouring lane by a cross-lane shuffle;
//   2. the horizontal pass as a max-plus scan: every lane summarises its K
//      columns as a transfer Y_out = M (x) Y_in (+) b, where Y is the
//      candidate value entering the lane's first column (H for linear, E
//      for affine, the pair (E, Q) for convex).  M is the same for every
//      lane (a scalar decay K*g or K*e, or a 2x2 max-plus matrix), so the
//      wavefront scan only needs its powers M^(2^t); waves are chained
//      through LDS;
//   3. a second pass over the K columns with the exact carry, the stores
//      of the H/F/E/O/Q rows and the end-cell bookkeeping.
// No sentinel ever takes part in the arithmetic: aggregates are formed
// with the carry term absent, so int32 scores may use their full range.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spoa_hip_dev {

constexpr int kMaxWaves = 16;

// Scoring parameters and scan constants for one launch.
struct KParams {
  int32_t type;     // 0 SW, 1 NW, 2 OV
  int32_t m, mis, g, e, q, c;
  int32_t pad;      // padding-column score
  int32_t ninf;     // the score type's minus infinity (row 0 of F and O)
  int32_t decay;    // linear: K*g; affine: K*e (per-lane decay of the scan)
  int32_t mpow[8][4];  // convex: M^(2^t) for t = 0..6 (row-major 2x2), [7] unused
};

// One (graph, read) job.  Offsets are in elements of the respective buffer.
struct DevJob {
  uint64_t mat_addr;   // device address of the job's DP matrices
  uint32_t graph_off;  // 32-bit words into the graph buffer
  uint32_t seq_off;    // bytes into the sequence buffer
  uint32_t V;          // graph nodes
  uint32_t n;          // read length
  uint32_t W;          // reference row width ceil(n / L) * L
  uint32_t stride;     // row stride of one matrix, a multiple of K
  uint32_t pair_off;   // pairs into the pair buffer
  uint32_t cap;        // pair capacity (V + n suffices)
};

// Graph block of a job (32-bit words from graph_off, 16-byte aligned):
//   rows[V] (uint4 per rank: decoded base | sink << 8 | predecessor count
//   << 9 (1 for a source, whose predecessor is row 0), first predecessor
//   row, offset of the predecessor list in preds, fcH of the row),
//   node_id[V], pred_off[V + 1] (relative to preds), fcH[V + 1],
//   fcF[V + 1], fcO[V + 1], preds[...] (predecessor rows, rank + 1, in
//   Node::inedges order).
struct GraphView {
  const uint4* rows;
  const int32_t* node_id;
  const uint32_t* pred_off;
  const int32_t* fcH;
  const int32_t* fcF;
  const int32_t* fcO;
  const uint32_t* preds;
};

__device__ __forceinline__ GraphView MakeView(const uint32_t* base, uint32_t V) {
  GraphView g;
  g.rows = reinterpret_cast<const uint4*>(base);
  g.node_id = reinterpret_cast<const int32_t*>(base + 4 * V);
  g.pred_off = base + 5 * V;
  g.fcH = reinterpret_cast<const int32_t*>(base + 6 * V + 1);
  g.fcF = g.fcH + (V + 1);
  g.fcO = g.fcF + (V + 1);
  g.preds = base + 9 * V + 4;
  return g;
}

struct KArgs {
  KParams prm;
  const DevJob* jobs;
  uint32_t n_jobs;
  uint32_t* queue;           // job counter, zero at launch
  const uint32_t* graphs;
  const uint8_t* seqs;
  int32_t* pairs;            // (node id, column) pairs in traceback order
  int32_t* results;          // per job: n_pairs, score, has_score, status
};

enum DevStatus : int32_t { kDevOk = 0, kDevCapacity = 6, kDevInconsistent = 7 };

// ------------------------------------------------------------- helpers --

template <typename ST, int K>
__device__ __forceinline__ void LoadCols(const ST* __restrict__ p, int32_t (&v)[K]) {
  if constexpr (sizeof(ST) == 2) {
    static_assert(K % 8 == 0, "int16 rows move 8 columns per 16-byte access");
    const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int i = 0; i < K / 8; ++i) {
      const uint4 u = q[i];
      const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[i * 8 + 2 * j] = static_cast<int32_t>(static_cast<int16_t>(w[j] & 0xffffu));
        v[i * 8 + 2 * j + 1] = static_cast<int32_t>(w[j]) >> 16;
      }
    }
  } else {
    static_assert(K % 4 == 0, "int32 rows move 4 columns per 16-byte access");
    const int4* q = reinterpret_cast<const int4*>(p);
#pragma unroll
    for (int i = 0; i < K / 4; ++i) {
      const int4 u = q[i];
      v[i * 4 + 0] = u.x;
      v[i * 4 + 1] = u.y;
      v[i * 4 + 2] = u.z;
      v[i * 4 + 3] = u.w;
    }
  }
}

template <typename ST, int K>
__device__ __forceinline__ void StoreCols(ST* __restrict__ p, const int32_t (&v)[K]) {
  if constexpr (sizeof(ST) == 2) {
    uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
    for (int i = 0; i < K / 8; ++i) {
      uint32_t w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        w[j] = (static_cast<uint32_t>(v[i * 8 + 2 * j]) & 0xffffu) |
               (static_cast<uint32_t>(v[i * 8 + 2 * j + 1]) << 16);
      }
      q[i] = make_uint4(w[0], w[1], w[2], w[3]);
    }
  } else {
    int4* q = reinterpret_cast<int4*>(p);
#pragma unroll
    for (int i = 0; i < K / 4; ++i) {
      q[i] = make_int4(v[i * 4 + 0], v[i * 4 + 1], v[i * 4 + 2], v[i * 4 + 3]);
    }
  }
}

// y <- M (x) y for a 2x2 max-plus matrix (row-major).
__device__ __forceinline__ void MpApply(const int32_t* M, int32_t& y0, int32_t& y1) {
  const int32_t a = max(M[0] + y0, M[1] + y1);
  const int32_t b = max(M[2] + y0, M[3] + y1);
  y0 = a;
  y1 = b;
}

__device__ __forceinline__ uint32_t Uniform(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

// The K read bases at columns c0.. (4 per word) and how many of them lie
// inside the read; the sequence block is padded to the row stride.
template <int K>
__device__ __forceinline__ void LoadChars(const uint8_t* __restrict__ seq, uint32_t c0, uint32_t n, bool active,
                                          uint32_t (&scp)[K / 4], int32_t& nvalid) {
  if (active) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(seq + c0);
#pragma unroll
    for (int i = 0; i < K / 4; ++i) scp[i] = w[i];
  } else {
#pragma unroll
    for (int i = 0; i < K / 4; ++i) scp[i] = 0u;
  }
  const int32_t left = static_cast<int32_t>(n) - static_cast<int32_t>(c0);
  nvalid = left < 0 ? 0 : (left > K ? K : left);
}

// ------------------------------------------------------------ traceback --
// Serial walk of the reference traceback, run by one lane.  Pairs are
// emitted in traceback order (the host reverses them).
template <int SUB, typename ST>
__device__ void Traceback(const KParams& prm, const GraphView& gv, const uint8_t* __restrict__ seq,
                          uint32_t n, const ST* __restrict__ mat, uint64_t rstride, uint32_t stride,
                          int32_t max_i, int32_t max_j, int32_t* __restrict__ out, uint32_t cap,
                          uint32_t* n_out, int32_t* status_out) {
  const int32_t g = prm.g, e = prm.e, q = prm.q, c = prm.c;
  const int type = prm.type;
  uint32_t np = 0;
  int32_t status = kDevOk;
  auto emit = [&](int32_t a, int32_t b) {
    if (np < cap) {
      out[2 * np] = a;
      out[2 * np + 1] = b;
    } else {
      status = kDevCapacity;
    }
    ++np;
  };
  // matrix t of row r at column col (t: 0 H, 1 F, 2 E, 3 O, 4 Q)
  auto at = [&](int t, int32_t r, int32_t col) -> int32_t {
    return static_cast<int32_t>(
        mat[static_cast<uint64_t>(r) * rstride + static_cast<uint64_t>(t) * stride + static_cast<uint32_t>(col)]);
  };
  auto prof = [&](int32_t r, int32_t col) -> int32_t {
    if (col >= static_cast<int32_t>(n)) return prm.pad;
    const uint32_t ch = gv.rows[r - 1].x & 0xffu;
    return seq[col] == ch ? prm.m : prm.mis;
  };

  int32_t i = max_i, j = max_j, prev_i = 0, prev_j = 0;
  while (true) {
    // every step emits one pair, so a full buffer also bounds the walk
    if (j == -1 || i == 0 || status != kDevOk) break;
    const int32_t h = at(0, i, j);
    if (type == 0 && h == 0) break;
    const uint32_t pb = gv.pred_off[i - 1], pe = gv.pred_off[i];
    const bool source = pb == pe;
    const uint32_t np_ = source ? 1u : pe - pb;
    auto pred = [&](uint32_t k) -> int32_t { return source ? 0 : static_cast<int32_t>(gv.preds[pb + k]); };
    bool found = false, ext_left = false, ext_up = false;
    // diagonal (:1053-1063)
    const int32_t pr = prof(i, j);
    for (uint32_t k = 0; k < np_; ++k) {
      const int32_t p = pred(k);
      const int32_t hd = j == 0 ? gv.fcH[p] : at(0, p, j - 1);
      if (h == hd + pr) {
        prev_i = p;
        prev_j = j - 1;
        found = true;
        break;
      }
    }
    // up (:1065-1074, :1447-1457, :1963-1975)
    if (!found) {
      for (uint32_t k = 0; k < np_; ++k) {
        const int32_t p = pred(k);
        const int32_t hu = at(0, p, j);
        bool hit;
        if (SUB == 0) {
          hit = h == hu + g;
        } else if (SUB == 1) {
          hit = (ext_up = (h == at(1, p, j) + e)) || h == hu + g;
        } else {
          hit = (ext_up = (h == at(1, p, j) + e)) || h == hu + g || (ext_up = (h == at(3, p, j) + c)) ||
                h == hu + q;
        }
        if (hit) {
          prev_i = p;
          prev_j = j;
          found = true;
          break;
        }
      }
    }
    // left (:1076-1083, :1459-1468, :1977-1990); column 0 reads the first column
    if (!found) {
      const int32_t hl = j > 0 ? at(0, i, j - 1) : gv.fcH[i];
      bool hit;
      if (SUB == 0) {
        hit = h == hl + g;
      } else {
        const int32_t el = j > 0 ? at(2, i, j - 1) : gv.fcH[i];
        if (SUB == 1) {
          hit = (ext_left = (h == el + e)) || h == hl + g;
        } else {
          const int32_t ql = j > 0 ? at(4, i, j - 1) : gv.fcH[i];
          hit = (ext_left = (h == el + e)) || h == hl + g || (ext_left = (h == ql + c)) || h == hl + q;
        }
      }
      if (hit) {
        prev_i = i;
        prev_j = j - 1;
        found = true;
      }
    }
    emit(i == prev_i ? -1 : gv.node_id[i - 1], j == prev_j ? -1 : j);
    i = prev_i;
    j = prev_j;

    if (ext_left) {
      // :1484-1509, :2006-2039 -- convex stops when either chain breaks
      while (status == kDevOk) {
        emit(-1, j);
        --j;
        if (j == -1) break;
        if (at(2, i, j) + e != at(2, i, j + 1)) break;
        if (SUB == 2 && at(4, i, j) + c != at(4, i, j + 1)) break;
      }
    } else if (ext_up) {
      // :1510-1547 affine, :2040-2093 convex
      while (i != 0 && status == kDevOk) {
        const uint32_t qb = gv.pred_off[i - 1], qe = gv.pred_off[i];
        const int32_t fv = at(1, i, j);
        bool stop;
        prev_i = 0;
        if (SUB == 1) {
          stop = false;
          for (uint32_t k = qb; k < qe; ++k) {
            const int32_t p = static_cast<int32_t>(gv.preds[k]);
            if ((stop = (fv == at(0, p, j) + g)) || fv == at(1, p, j) + e) {
              prev_i = p;
              break;
            }
          }
        } else {
          const int32_t ov = at(3, i, j);
          stop = true;
          for (uint32_t k = qb; k < qe; ++k) {
            const int32_t p = static_cast<int32_t>(gv.preds[k]);
            if (fv == at(1, p, j) + e || ov == at(3, p, j) + c) {
              prev_i = p;
              stop = false;
              break;
            }
          }
          if (stop) {
            for (uint32_t k = qb; k < qe; ++k) {
              const int32_t p = static_cast<int32_t>(gv.preds[k]);
              if (fv == at(0, p, j) + g || ov == at(0, p, j) + q) {
                prev_i = p;
                break;
              }
            }
          }
        }
        emit(gv.node_id[i - 1], -1);
        i = prev_i;
        if (stop || i == 0) break;
      }
    }
  }

  // NW: finish along row 0 / column 0 (:1102-1124, :1552-1574, :2098-2121)
  if (type == 1) {
    while (i == 0 && j != -1 && status == kDevOk) {
      emit(-1, j);
      --j;
    }
    while (i != 0 && j == -1 && status == kDevOk) {
      emit(gv.node_id[i - 1], -1);
      const uint32_t qb = gv.pred_off[i - 1], qe = gv.pred_off[i];
      if (qb == qe) {
        i = 0;
        continue;
      }
      int32_t next = i;
      for (uint32_t k = qb; k < qe; ++k) {
        const int32_t p = static_cast<int32_t>(gv.preds[k]);
        bool hit;
        if (SUB == 0) {
          hit = gv.fcH[i] == gv.fcH[p] + g;
        } else if (SUB == 1) {
          hit = gv.fcH[i] == gv.fcH[p] + e;
        } else {
          hit = gv.fcF[i] == gv.fcF[p] + e || gv.fcO[i] == gv.fcO[p] + c;
        }
        if (hit) {
          next = p;
          break;
        }
      }
      if (next == i) {
        status = kDevInconsistent;
        break;
      }
      i = next;
    }
  }
  *n_out = np;
  *status_out = status;
}

// ------------------------------------------------------- packed rows --
// A lane's K columns of one stored row as the 32-bit words the 16-byte
// loads and stores move (int16: two columns a word).  The previous row is
// kept in this form in registers, bit-identical to what memory holds.
template <typename ST, int K>
constexpr int Words() {
  return sizeof(ST) == 2 ? K / 2 : K;
}

template <typename ST, int K>
__device__ __forceinline__ void Pack(const int32_t (&v)[K], uint32_t (&w)[Words<ST, K>()]) {
  if constexpr (sizeof(ST) == 2) {
#pragma unroll
    for (int i = 0; i < K / 2; ++i) {
      w[i] = (static_cast<uint32_t>(v[2 * i]) & 0xffffu) | (static_cast<uint32_t>(v[2 * i + 1]) << 16);
    }
  } else {
#pragma unroll
    for (int i = 0; i < K; ++i) w[i] = static_cast<uint32_t>(v[i]);
  }
}

template <typename ST, int K>
__device__ __forceinline__ void Unpack(const uint32_t (&w)[Words<ST, K>()], int32_t (&v)[K]) {
  if constexpr (sizeof(ST) == 2) {
#pragma unroll
    for (int i = 0; i < K / 2; ++i) {
      v[2 * i] = static_cast<int32_t>(static_cast<int16_t>(w[i] & 0xffffu));
      v[2 * i + 1] = static_cast<int32_t>(w[i]) >> 16;
    }
  } else {
#pragma unroll
    for (int i = 0; i < K; ++i) v[i] = static_cast<int32_t>(w[i]);
  }
}

template <typename ST, int K>
__device__ __forceinline__ void LoadWords(const ST* __restrict__ p, uint32_t (&w)[Words<ST, K>()]) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
  for (int i = 0; i < Words<ST, K>() / 4; ++i) {
    const uint4 u = q[i];
    w[4 * i + 0] = u.x;
    w[4 * i + 1] = u.y;
    w[4 * i + 2] = u.z;
    w[4 * i + 3] = u.w;
  }
}

template <typename ST, int K>
__device__ __forceinline__ void StoreWords(ST* __restrict__ p, const uint32_t (&w)[Words<ST, K>()]) {
  uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
  for (int i = 0; i < Words<ST, K>() / 4; ++i) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// --------------------------------------------------------------- kernel --
// SUB: 0 linear, 1 affine, 2 convex.  ST: stored score type.  K: columns
// per lane.  MAXT: launch bound (threads per workgroup).
template <int SUB, typename ST, int K, int MAXT>
__global__ void __launch_bounds__(MAXT) PoaAlignKernel(const KArgs a) {
  constexpr int L = sizeof(ST) == 2 ? 16 : 8;
  constexpr int NM = SUB == 0 ? 1 : (SUB == 1 ? 3 : 5);  // H | F E | O Q
  constexpr int KW = Words<ST, K>();
  static_assert(SUB != 2 || K % L == 0, "convex lanes must start reference segments");
  static_assert(KW % 4 == 0, "rows move in 16-byte accesses");
  constexpr int32_t kMinI = -2147483647 - 1;

  __shared__ int32_t s_tot[2][kMaxWaves][2];
  __shared__ int32_t s_red[kMaxWaves][3];
  __shared__ uint32_t s_job;

  const KParams& prm = a.prm;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwaves = static_cast<int>(blockDim.x) >> 6;
  const int type = prm.type;
  const int32_t m = prm.m, mis = prm.mis, g = prm.g, e = prm.e, q = prm.q, cc = prm.c;
  const int32_t pad = prm.pad;

  for (;;) {
    if (tid == 0) s_job = atomicAdd(a.queue, 1u);
    __syncthreads();
    const uint32_t jx = Uniform(s_job);
    __syncthreads();
    if (jx >= a.n_jobs) break;

    const DevJob J = a.jobs[jx];
    const uint32_t V = Uniform(J.V), n = Uniform(J.n), W = Uniform(J.W), stride = Uniform(J.stride);
    const GraphView gv = MakeView(a.graphs + J.graph_off, V);
    const uint8_t* __restrict__ seq = a.seqs + J.seq_off;
    ST* __restrict__ mat = reinterpret_cast<ST*>(J.mat_addr);
    const uint64_t rstride = static_cast<uint64_t>(stride) * NM;
    const uint32_t tile_w = static_cast<uint32_t>(nwaves) * 64u * K;
    const uint32_t ntiles = (stride + tile_w - 1) / tile_w;
    const bool single = ntiles == 1;  // a row is one tile: the previous row stays in registers
    const bool sync = nwaves > 1 || ntiles > 1;

    // row 0 (:556-694): E = g + c*e, Q = q + c*c, F = O = -inf, H per type
    auto row0_h = [&](int32_t ci) -> int32_t {
      if (type == 0) return 0;
      if (SUB == 0) return g + ci * g;
      if (SUB == 1) return g + ci * e;
      return max(g + ci * e, q + ci * cc);
    };
    for (uint32_t col = tid; col < stride; col += blockDim.x) {
      const int32_t ci = static_cast<int32_t>(col);
      mat[col] = static_cast<ST>(row0_h(ci));
      if (SUB >= 1) {
        mat[stride + col] = static_cast<ST>(prm.ninf);
        mat[2 * stride + col] = static_cast<ST>(g + ci * e);
      }
      if (SUB == 2) {
        mat[3 * stride + col] = static_cast<ST>(prm.ninf);
        mat[4 * stride + col] = static_cast<ST>(q + ci * cc);
      }
    }
    __syncthreads();

    // End-cell candidates per lane: the lane's best (floored) value and the
    // first row reaching it.  SW starts at 0, OV and NW at the type's -inf.
    int32_t best = type == 0 ? 0 : prm.ninf;
    int32_t best_row = -1;
    int par = 0;
    // the read's bases under this lane's columns (reloaded per tile when a
    // row spans several tiles)
    uint32_t scp[K / 4];
    int32_t nvalid;
    LoadChars<K>(seq, static_cast<uint32_t>(tid) * K, n, static_cast<uint32_t>(tid) * K < stride, scp, nvalid);
    // the previous row (H, F, O) of this lane's columns, starting with row 0
    uint32_t cH[KW], cF[KW], cO[KW];
    if (single) {
      int32_t v[K], f[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        v[k] = row0_h(tid * K + k);
        f[k] = prm.ninf;
      }
      Pack<ST, K>(v, cH);
      Pack<ST, K>(f, cF);
      Pack<ST, K>(f, cO);
    }
    int32_t fch_prev = gv.fcH[0];
    uint4 rec = gv.rows[0];

    for (uint32_t r = 1; r <= V; ++r) {
      // row record: base | sink | predecessor count, first predecessor row,
      // offset of the predecessor list, first-column score (next one prefetched)
      const uint4 cur = rec;
      if (r < V) rec = gv.rows[r];
      const uint32_t rch = cur.x & 0xffu;
      const bool sink = ((cur.x >> 8) & 1u) != 0;
      const uint32_t npred = cur.x >> 9;
      const uint32_t p0 = cur.y, pb = cur.z;
      const int32_t fch_r = static_cast<int32_t>(cur.w);
      ST* __restrict__ row = mat + static_cast<uint64_t>(r) * rstride;
      // carry entering column 0: linear H, affine E = fc + g, convex (E, Q)
      int32_t y0 = fch_r + g;
      int32_t y1 = fch_r + q;

      for (uint32_t t = 0; t < ntiles; ++t) {
        const uint32_t c0 = t * tile_w + static_cast<uint32_t>(tid) * K;
        const bool active = c0 < stride;

        // ---- 1. vertical pass over the predecessors, in in-edge order
        if (!single) LoadChars<K>(seq, c0, n, active, scp, nvalid);
        int32_t h1[K], fv[K], ov[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          h1[k] = kMinI;
          fv[k] = kMinI;
          ov[k] = kMinI;
        }
        for (uint32_t pi = 0; pi < npred; ++pi) {
          const uint32_t p = pi == 0 ? p0 : Uniform(gv.preds[pb + pi]);
          const ST* __restrict__ prow = mat + static_cast<uint64_t>(p) * rstride;
          const bool cached = single && p + 1 == r;
          int32_t hv[K];
          if (cached) {
            Unpack<ST, K>(cH, hv);
          } else if (active) {
            LoadCols<ST, K>(prow + c0, hv);
          } else {
#pragma unroll
            for (int k = 0; k < K; ++k) hv[k] = 0;
          }
          // H_p[c0 - 1]: neighbouring lane, or the first column / memory
          // across waves and tiles
          int32_t hm1 = __shfl_up(hv[K - 1], 1, 64);
          if (lane == 0) {
            if (c0 == 0) {
              hm1 = p + 1 == r ? fch_prev : gv.fcH[p];
            } else {
              hm1 = active ? static_cast<int32_t>(prow[c0 - 1]) : 0;
            }
          }
          // diagonal (:838-850): H_p[c - 1] + profile
          {
            int32_t prev = hm1;
#pragma unroll
            for (int k = 0; k < K; ++k) {
              const uint32_t ch = (scp[k / 4] >> (8 * (k % 4))) & 0xffu;
              const int32_t pr = k < nvalid ? (ch == rch ? m : mis) : pad;
              h1[k] = max(h1[k], prev + pr);
              prev = hv[k];
            }
          }
          if (SUB == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) fv[k] = max(fv[k], hv[k] + g);
          } else {
            int32_t fp[K];
            if (cached) {
              Unpack<ST, K>(cF, fp);
            } else if (active) {
              LoadCols<ST, K>(prow + stride + c0, fp);
            } else {
#pragma unroll
              for (int k = 0; k < K; ++k) fp[k] = 0;
            }
#pragma unroll
            for (int k = 0; k < K; ++k) fv[k] = max(fv[k], max(hv[k] + g, fp[k] + e));
            if (SUB == 2) {
              int32_t op[K];
              if (cached) {
                Unpack<ST, K>(cO, op);
              } else if (active) {
                LoadCols<ST, K>(prow + 3 * stride + c0, op);
              } else {
#pragma unroll
                for (int k = 0; k < K; ++k) op[k] = 0;
              }
#pragma unroll
              for (int k = 0; k < K; ++k) ov[k] = max(ov[k], max(hv[k] + q, op[k] + cc));
            }
          }
        }
        // H1: the best of the vertical and diagonal moves
#pragma unroll
        for (int k = 0; k < K; ++k) {
          h1[k] = max(h1[k], fv[k]);
          if (SUB == 2) h1[k] = max(h1[k], ov[k]);
        }

        // ---- 2. lane transfer with the carry absent
        int32_t b0 = 0, b1 = 0;
        if (SUB == 0) {
          int32_t x = h1[0];
#pragma unroll
          for (int k = 1; k < K; ++k) x = max(h1[k], x + g);
          b0 = x + g;
        } else if (SUB == 1) {
          int32_t x = h1[0] + g;  // E at column c0 + 1
#pragma unroll
          for (int k = 2; k < K; ++k) x = max(h1[k - 1] + g, x + e);
          b0 = max(h1[K - 1] + g, x + e);
        } else {
          int32_t xe = h1[0] + g, xq = h1[0] + q;  // E, Q at column c0 + 1
#pragma unroll
          for (int k = 2; k < K; ++k) {
            if (k % L == 0) {
              const int32_t hf = max(h1[k - 1], max(xe, xq));
              const int32_t ne = max(hf + g, xe + e);
              const int32_t nq = max(hf + q, xq + cc);
              xe = ne;
              xq = nq;
            } else {
              xe = max(h1[k - 1] + g, xe + e);
              xq = max(h1[k - 1] + q, xq + cc);
            }
          }
          const int32_t hf = max(h1[K - 1], max(xe, xq));
          b0 = max(hf + g, xe + e);
          b1 = max(hf + q, xq + cc);
        }

        // ---- wavefront inclusive scan of the transfers
        int32_t s0 = b0, s1 = b1;
        if (SUB != 2) {
          const int32_t d = prm.decay;
#pragma unroll
          for (int st = 0; st < 6; ++st) {
            const int sh = 1 << st;
            const int32_t y = __shfl_up(s0, sh, 64);
            if (lane >= sh) s0 = max(s0, y + sh * d);
          }
        } else {
#pragma unroll
          for (int st = 0; st < 6; ++st) {
            const int sh = 1 << st;
            int32_t z0 = __shfl_up(s0, sh, 64);
            int32_t z1 = __shfl_up(s1, sh, 64);
            if (lane >= sh) {
              MpApply(prm.mpow[st], z0, z1);
              s0 = max(s0, z0);
              s1 = max(s1, z1);
            }
          }
        }
        // carry into this wave: the tile carry folded with earlier waves
        int32_t w0 = y0, w1 = y1;
        int32_t tile_out0, tile_out1;
        if (nwaves > 1) {
          if (lane == 63) {
            s_tot[par][wave][0] = s0;
            s_tot[par][wave][1] = s1;
          }
          __syncthreads();
          int32_t f0 = y0, f1 = y1;
          for (int w = 0; w < nwaves; ++w) {
            if (w == wave) {
              w0 = f0;
              w1 = f1;
            }
            const int32_t t0 = s_tot[par][w][0], t1 = s_tot[par][w][1];
            if (SUB != 2) {
              f0 = max(t0, f0 + 64 * prm.decay);
            } else {
              MpApply(prm.mpow[6], f0, f1);
              f0 = max(f0, t0);
              f1 = max(f1, t1);
            }
          }
          tile_out0 = f0;
          tile_out1 = f1;
          par ^= 1;
        } else {
          const int32_t t0 = __shfl(s0, 63, 64), t1 = __shfl(s1, 63, 64);
          int32_t f0 = y0, f1 = y1;
          if (SUB != 2) {
            f0 = max(t0, f0 + 64 * prm.decay);
          } else {
            MpApply(prm.mpow[6], f0, f1);
            f0 = max(f0, t0);
            f1 = max(f1, t1);
          }
          tile_out0 = f0;
          tile_out1 = f1;
        }
        // exclusive carry into this lane
        const int32_t e0 = __shfl_up(s0, 1, 64);
        const int32_t e1 = __shfl_up(s1, 1, 64);
        int32_t x0, x1;
        if (SUB != 2) {
          x0 = lane == 0 ? w0 : max(e0, w0 + lane * prm.decay);
          x1 = 0;
        } else {
          int32_t u0 = w0, u1 = w1;
#pragma unroll
          for (int st = 0; st < 6; ++st) {
            if ((lane >> st) & 1) MpApply(prm.mpow[st], u0, u1);
          }
          x0 = lane == 0 ? w0 : max(e0, u0);
          x1 = lane == 0 ? w1 : max(e1, u1);
        }

        // ---- 3. exact pass, stores, end-cell bookkeeping
        if (active) {
          int32_t hs[K];
          if (SUB == 0) {
            int32_t x = max(h1[0], x0);
            hs[0] = type == 0 ? max(x, 0) : x;
#pragma unroll
            for (int k = 1; k < K; ++k) {
              x = max(h1[k], x + g);
              hs[k] = type == 0 ? max(x, 0) : x;
            }
          } else {
            int32_t ev[K], qv[K];
            ev[0] = x0;
            qv[0] = x1;
#pragma unroll
            for (int k = 1; k < K; ++k) {
              if (SUB == 2 && k % L == 0) {
                const int32_t hf = max(h1[k - 1], max(ev[k - 1], qv[k - 1]));
                ev[k] = max(hf + g, ev[k - 1] + e);
                qv[k] = max(hf + q, qv[k - 1] + cc);
              } else {
                ev[k] = max(h1[k - 1] + g, ev[k - 1] + e);
                if (SUB == 2) qv[k] = max(h1[k - 1] + q, qv[k - 1] + cc);
              }
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
              int32_t h = max(h1[k], ev[k]);
              if (SUB == 2) h = max(h, qv[k]);
              hs[k] = type == 0 ? max(h, 0) : h;
            }
            uint32_t wF[KW], wE[KW];
            Pack<ST, K>(fv, wF);
            Pack<ST, K>(ev, wE);
            StoreWords<ST, K>(row + stride + c0, wF);
            StoreWords<ST, K>(row + 2 * stride + c0, wE);
            if (single) {
#pragma unroll
              for (int i = 0; i < KW; ++i) cF[i] = wF[i];
            }
            if (SUB == 2) {
              uint32_t wO[KW], wQ[KW];
              Pack<ST, K>(ov, wO);
              Pack<ST, K>(qv, wQ);
              StoreWords<ST, K>(row + 3 * stride + c0, wO);
              StoreWords<ST, K>(row + 4 * stride + c0, wQ);
              if (single) {
#pragma unroll
                for (int i = 0; i < KW; ++i) cO[i] = wO[i];
              }
            }
          }
          uint32_t wH[KW];
          Pack<ST, K>(hs, wH);
          StoreWords<ST, K>(row + c0, wH);
          if (single) {
#pragma unroll
            for (int i = 0; i < KW; ++i) cH[i] = wH[i];
          }

          // end-cell candidates (:903-928): SW all rows, OV sink rows (row
          // maxima floored at 0, :252-262), NW sink rows at column n - 1
          if (type == 0 || (type == 2 && sink)) {
            int32_t lm = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
              if (c0 + k < W) lm = max(lm, hs[k]);
            }
            if (best < lm) {
              best = lm;
              best_row = static_cast<int32_t>(r);
            }
          } else if (type == 1 && sink) {
            const uint32_t cn = n - 1;
            if (cn >= c0 && cn < c0 + K) {
              int32_t v = hs[0];
#pragma unroll
              for (int k = 1; k < K; ++k) {
                if (c0 + k == cn) v = hs[k];
              }
              if (best < v) {
                best = v;
                best_row = static_cast<int32_t>(r);
              }
            }
          }
        }
        y0 = tile_out0;
        y1 = tile_out1;
        if (sync) __syncthreads();  // the next tile / row reads these stores
      }
      fch_prev = fch_r;
    }
    // the traceback reads what every wave stored
    __threadfence_block();

    // ---- end cell: the best value, then the first row reaching it
    {
      int32_t v = best, rr = best_row < 0 ? 0x7fffffff : best_row;
      // lanes without a candidate row do not compete
      if (best_row < 0) v = kMinI;
#pragma unroll
      for (int st = 32; st >= 1; st >>= 1) {
        const int32_t ov2 = __shfl_xor(v, st, 64);
        const int32_t or2 = __shfl_xor(rr, st, 64);
        if (ov2 > v || (ov2 == v && or2 < rr)) {
          v = ov2;
          rr = or2;
        }
      }
      if (lane == 0) {
        s_red[wave][0] = v;
        s_red[wave][1] = rr;
      }
    }
    __syncthreads();
    int32_t max_score = kMinI, max_i = 0x7fffffff;
    for (int w = 0; w < nwaves; ++w) {
      const int32_t v = s_red[w][0], rr = s_red[w][1];
      if (v > max_score || (v == max_score && rr < max_i)) {
        max_score = v;
        max_i = rr;
      }
    }
    const bool found = max_i != 0x7fffffff;
    __syncthreads();

    // max_j (:937-953): NW the last column; SW/OV the first column of row
    // max_i (padding included) holding the best value, -1 if none
    int32_t max_j = static_cast<int32_t>(n) - 1;
    if (found && type != 1) {
      const ST* __restrict__ hrow = mat + static_cast<uint64_t>(max_i) * rstride;
      int32_t first = 0x7fffffff;
      for (uint32_t col = tid; col < W; col += blockDim.x) {
        if (static_cast<int32_t>(hrow[col]) == max_score) {
          first = static_cast<int32_t>(col);
          break;
        }
      }
#pragma unroll
      for (int st = 32; st >= 1; st >>= 1) first = min(first, __shfl_xor(first, st, 64));
      if (lane == 0) s_red[wave][2] = first;
      __syncthreads();
      first = 0x7fffffff;
      for (int w = 0; w < nwaves; ++w) first = min(first, s_red[w][2]);
      max_j = first == 0x7fffffff ? -1 : first;
      __syncthreads();
    }

    // ---- traceback in one lane
    if (tid == 0) {
      int32_t* res = a.results + 4 * jx;
      if (!found) {
        res[0] = 0;
        res[1] = 0;
        res[2] = 0;
        res[3] = kDevOk;
      } else {
        uint32_t np = 0;
        int32_t status = kDevOk;
        Traceback<SUB, ST>(prm, gv, seq, n, mat, rstride, stride, max_i, max_j, a.pairs + 2ull * J.pair_off,
                           J.cap, &np, &status);
        res[0] = static_cast<int32_t>(np);
        res[1] = max_score;
        res[2] = 1;
        res[3] = status;
      }
    }
    __syncthreads();
  }
}

// Zeroes the job counters of a pass (a kernel rather than a memset so the
// launch sequence can be captured in a graph).
__global__ void ResetCounters(uint32_t* p, uint32_t n) {
  if (blockIdx.x == 0 && threadIdx.x < n) p[threadIdx.x] = 0u;
}

}  // namespace spoa_hip_dev
