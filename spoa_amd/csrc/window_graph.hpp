// Device-resident POA window graphs: the reference Graph's AddAlignment and
// TopologicalSort (src/graph.cpp:76-303) run by one wavefront per window,
// and the rank-ordered graph block the alignment kernels read (the layout of
// MakeView in poa_kernels.hpp) built straight from the result.
//
// Written once for a "team" of lanes: on the device a team is one wave of
// 64 lanes (WaveTeam, in resident.hip), in the CPU unit test a single lane
// that runs every loop serially.  Every loop over nodes, edges or pairs is a
// sequence of team-wide chunks whose lanes write disjoint locations, and
// every ordered quantity (new node ids, new edge ids) comes from a prefix
// count over the chunks, so one lane and 64 lanes build the same graph.
//
// A window's graph lives in one region of device memory:
//   header (sizes, status, the character coder / decoder),
//   nodes:  code, in-degree, out-degree, first and last out-edge,
//           aligned-node count and list (at most kMaxAligned, in the order of
//           Node::aligned_nodes), rank_of, rank_to_node,
//   edges:  tail, head, weight (int64), position in the head's in-edge list,
//           next out-edge of the tail (out-edge lists are creation-ordered
//           chains),
//   sequences: every read's path of node ids (a read visits a node at most
//           once, so its path is what following its label from its first
//           node visits, src/graph.cpp:22-30), the strand it was added on,
//   scratch for the topological sort when the window does not fit in LDS.
//
// Parallel AddAlignment.  The reference adds an alignment's nodes and edges
// one pair at a time (src/graph.cpp:196-244).  Node ids and edge ids are
// handed out in that order, so their values are prefix counts.  A path
// visits a node at most once and never visits two nodes of one aligned group
// (no graph path joins two aligned nodes: a group's members are created on
// different reads' paths at the same column, and an alignment only moves
// along existing edges), so every node gains at most one in-edge and one
// out-edge per alignment and every group at most one member: the pairs can be
// merged in parallel.  A path that did reach two nodes of one group (ranks
// within one group's contiguous block) is reported as an error, never merged.
#pragma once

#include <cstdint>

namespace spoa_amd {
namespace wg {

constexpr uint32_t kNil = 0xffffffffu;
constexpr uint32_t kMaxAligned = 7;  // aligned groups of up to 8 codes
constexpr uint16_t kNoCode = 0xffffu;
constexpr uint32_t kSortSink = 64;  // topological sort: a slot per lane for writes nothing reads

enum WinStatus : int32_t {
  kWinOk = 0,
  kWinAlignFailed = 1,       // the alignment job of the round failed
  kWinInvalidAlignment = 2,  // "[spoa::Graph::AddAlignment] error: invalid alignment"
  kWinMissingSequence = 3,   // "[spoa::Graph::AddAlignment] error: missing sequence in alignment"
  kWinGroupTooLarge = 4,     // an aligned group of more than kMaxAligned + 1 codes
  kWinGroupRevisit = 5,      // an alignment through two nodes of one aligned group
  kWinCapacity = 6,          // more nodes, edges or path entries than the region holds
};

struct WinHeader {
  uint32_t V, E, nseq, ncodes;
  uint32_t path_len;
  uint32_t R;  // sum over the nodes of max(1, in-degree): the rows the DP reads
  uint32_t nslots, slot_chars;
  int32_t status;
  uint32_t pad[7];
  uint16_t coder[256];  // code of a character, kNoCode if unseen
  uint8_t decoder[256];
};
static_assert(sizeof(WinHeader) == 832, "header layout");

// Capacities of one window's region.
struct WinCaps {
  uint32_t V;  // nodes (the window's total read length bounds them)
  uint32_t E;  // edges (likewise)
  uint32_t L;  // path entries (the total read length)
  uint32_t S;  // sequences (reads)
};

__host__ __device__ inline uint64_t Align256(uint64_t x) { return (x + 255) / 256 * 256; }

// Byte offsets of the arrays of one window's region.
struct WinLayout {
  uint64_t hdr, code, indeg, outdeg, out_head, out_tail, al_cnt, al, rank_of, rank_to_node;
  uint64_t e_tail, e_head, e_in_pos, e_next_out, e_weight;
  uint64_t path, path_off, reversed;
  uint64_t scratch;  // topological sort scratch (global mode), 32-bit words
  uint64_t scratch_words;
  uint64_t total;
};

// Scratch words the sort needs in global mode (32-bit indices): in-edge
// offsets and tails by node id, aligned-list offsets and entries, the stack
// (every push follows an in-edge or an aligned entry of a node visited once,
// or starts a root), for global alignments the source distance of every
// node, a sink, then the marks (a byte each) at an 8-byte boundary and
// their sink.
__host__ __device__ inline uint64_t SortScratchWords(uint64_t V, uint64_t E) {
  const uint64_t A = V * kMaxAligned;
  return (V + 1) + E + (V + 1) + A + (V + E + A + 1) + V + kSortSink + 1 + (V + 7) / 8 * 2 + kSortSink / 4;
}

__host__ __device__ inline WinLayout MakeLayout(const WinCaps& c) {
  WinLayout l{};
  uint64_t o = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t at = o;
    o = Align256(o + bytes);
    return at;
  };
  l.hdr = take(sizeof(WinHeader));
  l.code = take(c.V);
  l.indeg = take(4ull * c.V);
  l.outdeg = take(4ull * c.V);
  l.out_head = take(4ull * c.V);
  l.out_tail = take(4ull * c.V);
  l.al_cnt = take(c.V);
  l.al = take(4ull * c.V * kMaxAligned);
  l.rank_of = take(4ull * c.V);
  l.rank_to_node = take(4ull * c.V);
  l.e_tail = take(4ull * c.E);
  l.e_head = take(4ull * c.E);
  l.e_in_pos = take(4ull * c.E);
  l.e_next_out = take(4ull * c.E);
  l.e_weight = take(8ull * c.E);
  l.path = take(4ull * c.L);
  l.path_off = take(4ull * (c.S + 1));
  l.reversed = take(c.S);
  l.scratch_words = SortScratchWords(c.V, c.E);
  l.scratch = take(4ull * l.scratch_words);
  l.total = o;
  return l;
}

struct WinView {
  WinHeader* hdr;
  uint8_t* code;
  uint32_t *indeg, *outdeg, *out_head, *out_tail;
  uint8_t* al_cnt;
  uint32_t* al;
  uint32_t *rank_of, *rank_to_node;
  uint32_t *e_tail, *e_head, *e_in_pos, *e_next_out;
  int64_t* e_weight;
  uint32_t *path, *path_off;
  uint8_t* reversed;
  uint32_t* scratch;
  WinCaps caps;
};

__host__ __device__ inline WinView MakeWinView(uint8_t* base, const WinCaps& c) {
  const WinLayout l = MakeLayout(c);
  WinView v;
  v.hdr = reinterpret_cast<WinHeader*>(base + l.hdr);
  v.code = base + l.code;
  v.indeg = reinterpret_cast<uint32_t*>(base + l.indeg);
  v.outdeg = reinterpret_cast<uint32_t*>(base + l.outdeg);
  v.out_head = reinterpret_cast<uint32_t*>(base + l.out_head);
  v.out_tail = reinterpret_cast<uint32_t*>(base + l.out_tail);
  v.al_cnt = base + l.al_cnt;
  v.al = reinterpret_cast<uint32_t*>(base + l.al);
  v.rank_of = reinterpret_cast<uint32_t*>(base + l.rank_of);
  v.rank_to_node = reinterpret_cast<uint32_t*>(base + l.rank_to_node);
  v.e_tail = reinterpret_cast<uint32_t*>(base + l.e_tail);
  v.e_head = reinterpret_cast<uint32_t*>(base + l.e_head);
  v.e_in_pos = reinterpret_cast<uint32_t*>(base + l.e_in_pos);
  v.e_next_out = reinterpret_cast<uint32_t*>(base + l.e_next_out);
  v.e_weight = reinterpret_cast<int64_t*>(base + l.e_weight);
  v.path = reinterpret_cast<uint32_t*>(base + l.path);
  v.path_off = reinterpret_cast<uint32_t*>(base + l.path_off);
  v.reversed = base + l.reversed;
  v.scratch = reinterpret_cast<uint32_t*>(base + l.scratch);
  v.caps = c;
  return v;
}

// Graph-block words of a window with V nodes and E edges (MakeView layout).
__host__ __device__ inline uint64_t BlockWords(uint64_t V, uint64_t E) { return (9 * V + 4 + E + 3) / 4 * 4; }

// An empty window: no nodes, no codes.
template <class Team>
__host__ __device__ inline void InitWindow(const Team& t, const WinView& g) {
  for (uint32_t i = t.lane; i < 256; i += t.size) {
    g.hdr->coder[i] = kNoCode;
    g.hdr->decoder[i] = 0;
  }
  if (t.lane == 0) {
    g.hdr->V = g.hdr->E = g.hdr->nseq = g.hdr->ncodes = 0;
    g.hdr->path_len = 0;
    g.hdr->R = 0;
    g.hdr->nslots = g.hdr->slot_chars = 0;
    g.hdr->status = kWinOk;
    g.path_off[0] = 0;
  }
  t.sync();
}

// One read and the alignment to add (pairs as the alignment kernel stores
// them: np (node id, column) pairs in traceback order, i.e. path order
// reversed).
struct AlignIn {
  const int32_t* pairs;
  uint32_t np;
  const uint8_t* seq;
  uint32_t len;
  const uint8_t* qual;  // Phred+33 per base, or null (weight 1)
  bool qual_reversed;   // the read is the reverse complement: weight i is qual[len - 1 - i]
};

__host__ __device__ inline uint32_t WeightAt(const AlignIn& in, uint32_t i) {
  if (!in.qual) return 1u;
  const uint32_t k = in.qual_reversed ? in.len - 1 - i : i;
  return static_cast<uint32_t>(in.qual[k]) - 33u;
}

// Appends edge e = tail -> head (new) to the tail's out-list and the head's
// in-list.  Each node gains at most one in-edge and one out-edge per
// alignment, so concurrent calls touch disjoint nodes.
__host__ __device__ inline void NewEdge(const WinView& g, uint32_t e, uint32_t tail, uint32_t head, int64_t w) {
  g.e_tail[e] = tail;
  g.e_head[e] = head;
  g.e_weight[e] = w;
  g.e_next_out[e] = kNil;
  g.e_in_pos[e] = g.indeg[head];
  g.indeg[head] += 1;
  const uint32_t last = g.out_tail[tail];
  if (last == kNil) {
    g.out_head[tail] = e;
  } else {
    g.e_next_out[last] = e;
  }
  g.out_tail[tail] = e;
  g.outdeg[tail] += 1;
}

__host__ __device__ inline void NewNode(const WinView& g, uint32_t v, uint8_t code) {
  g.code[v] = code;
  g.indeg[v] = 0;
  g.outdeg[v] = 0;
  g.out_head[v] = kNil;
  g.out_tail[v] = kNil;
  g.al_cnt[v] = 0;
}

// Graph::AddAlignment (src/graph.cpp:155-247) up to the topological sort.
// Returns a WinStatus; on an error the graph is left as it was before the
// call where the error is found before any change (every check but the
// capacity of an aligned group precedes the first write).
template <class Team>
__host__ __device__ int32_t AddAlignment(const Team& t, const WinView& g, const AlignIn& in) {
  WinHeader& h = *g.hdr;
  const uint32_t len = in.len;
  if (len == 0) return kWinOk;  // :163-165
  // :172-177 -- codes in order of first appearance
  {
    bool unseen = false;
    for (uint32_t i = t.lane; i < len; i += t.size) unseen |= h.coder[in.seq[i]] == kNoCode;
    if (t.any(unseen)) {
      if (t.lane == 0) {
        for (uint32_t i = 0; i < len; ++i) {
          const uint8_t ch = in.seq[i];
          if (h.coder[ch] == kNoCode) {
            h.coder[ch] = static_cast<uint16_t>(h.ncodes);
            h.decoder[h.ncodes] = ch;
            h.ncodes += 1;
          }
        }
      }
      t.sync();
    }
  }
  const uint32_t V0 = h.V, E0 = h.E, P0 = h.path_len, s = h.nseq;
  if (static_cast<uint64_t>(V0) + len > g.caps.V || static_cast<uint64_t>(E0) + len > g.caps.E ||
      static_cast<uint64_t>(P0) + len > g.caps.L || s + 1 > g.caps.S) {
    return kWinCapacity;
  }
  const uint32_t np = in.np;
  auto pair_node = [&](uint32_t i) { return in.pairs[2ull * (np - 1 - i)]; };
  auto pair_pos = [&](uint32_t i) { return in.pairs[2ull * (np - 1 - i) + 1]; };

  if (np == 0) {
    // :179-183 -- the whole read as a new chain
    for (uint32_t i = t.lane; i < len; i += t.size) {
      const uint32_t v = V0 + i;
      NewNode(g, v, static_cast<uint8_t>(h.coder[in.seq[i]]));
      g.path[P0 + i] = v;
    }
    t.sync();
    for (uint32_t i = t.lane + 1; i < len; i += t.size) {
      NewEdge(g, E0 + i - 1, V0 + i - 1, V0 + i, static_cast<int64_t>(WeightAt(in, i - 1) + WeightAt(in, i)));
    }
    t.sync();
    if (t.lane == 0) {
      h.V = V0 + len;
      h.E = E0 + len - 1;
      h.path_len = P0 + len;
      g.path_off[s + 1] = P0 + len;
      h.nseq = s + 1;
    }
    t.sync();
    return kWinOk;
  }

  // :185-199 -- the first and last aligned read positions; positions must be
  // in range, and (alignments of the alignment kernels) consecutive
  uint32_t kfirst = 0xffffffffu, klast = 0, bad = 0;
  for (uint32_t base = 0; base < np; base += t.size) {
    const uint32_t i = base + t.lane;
    if (i < np) {
      const int32_t pos = pair_pos(i);
      if (pos != -1) {
        if (pos < 0 || pos >= static_cast<int32_t>(len)) bad = 1;
        kfirst = i < kfirst ? i : kfirst;
        klast = i > klast ? i : klast;
      }
    }
  }
  kfirst = t.min_all(kfirst);
  klast = t.max_all(klast);
  if (t.max_all(bad)) return kWinInvalidAlignment;
  if (kfirst == 0xffffffffu) return kWinMissingSequence;
  const uint32_t first_pos = static_cast<uint32_t>(pair_pos(kfirst));
  const uint32_t last_pos = static_cast<uint32_t>(pair_pos(klast));
  if (last_pos < first_pos) return kWinInvalidAlignment;
  const uint32_t nvalid = last_pos - first_pos + 1;
  const uint32_t npre = first_pos, nsuf = len - last_pos - 1;
  const uint32_t vnew0 = V0 + npre + nsuf;  // first node id of the aligned pairs' new nodes

  // Classify the aligned pairs in path order: the node each maps to
  // (:204-236) -- its own node, an aligned node of its code, or a new node
  // (joining the aligned group of the node it is aligned to).  The aligned
  // part of the path is written as it goes; the join targets of the new
  // nodes wait in the sort scratch.
  uint32_t* const join = g.scratch;
  uint32_t seen = 0, made = 0;
  uint32_t err = 0;
  for (uint32_t base = 0; base < np; base += t.size) {
    const uint32_t i = base + t.lane;
    const bool live = i < np;
    const int32_t pos = live ? pair_pos(i) : -1;
    const bool valid = pos != -1;
    const int32_t id = valid ? pair_node(i) : -1;
    uint32_t node = kNil, jt = kNil;
    bool fresh = false;
    uint32_t c = 0;
    if (valid) {
      c = h.coder[in.seq[pos]];
      if (id == -1) {
        fresh = true;
      } else {
        const uint32_t n = static_cast<uint32_t>(id);
        if (n >= V0) {
          err = kWinInvalidAlignment;
        } else if (g.code[n] == c) {
          node = n;
        } else {
          const uint32_t cnt = g.al_cnt[n];
          for (uint32_t k = 0; k < cnt; ++k) {
            const uint32_t a = g.al[n * kMaxAligned + k];
            if (g.code[a] == c) {
              node = a;
              break;
            }
          }
          if (node == kNil) {
            if (cnt >= kMaxAligned) err = kWinGroupTooLarge;
            fresh = true;
            jt = n;
          }
        }
      }
    }
    uint32_t vtot = 0, ftot = 0;
    const uint32_t vx = t.scan(valid ? 1u : 0u, &vtot);
    const uint32_t fx = t.scan(fresh ? 1u : 0u, &ftot);
    if (valid) {
      const uint32_t k = seen + vx;  // index of the pair among the aligned ones
      if (static_cast<uint32_t>(pos) != first_pos + k) err = kWinInvalidAlignment;
      if (fresh) {
        node = vnew0 + made + fx;
        join[made + fx] = jt;
      }
      if (k < nvalid) g.path[P0 + npre + k] = node;
    }
    seen += vtot;
    made += ftot;
  }
  t.sync();
  if (t.max_all(err)) return static_cast<int32_t>(t.max_all(err));
  if (seen != nvalid) return kWinInvalidAlignment;
  const uint32_t nfresh = made;

  // No two nodes of one aligned group on the path.  Every pair with a node
  // id is a row the alignment walked, ranks rise along the walk, and the
  // ranks of a group form a contiguous block led by its first-ranked
  // member, so a second row of a group would follow another row of the
  // block.  (Checked on the rows walked: a pair maps into its row's group.)
  {
    uint32_t hit = 0;
    for (uint32_t i = t.lane + 1; i < np; i += t.size) {
      const int32_t id = pair_node(i);
      if (id < 0) continue;
      const uint32_t v = static_cast<uint32_t>(id);
      const uint32_t cnt = g.al_cnt[v];
      if (cnt == 0) continue;
      uint32_t lead = g.rank_of[v];
      for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t r = g.rank_of[g.al[v * kMaxAligned + j]];
        lead = r < lead ? r : lead;
      }
      for (uint32_t j = i; j-- > 0;) {
        const int32_t u = pair_node(j);
        if (u >= 0) {
          if (g.rank_of[u] >= lead) hit = 1;
          break;
        }
      }
    }
    if (t.max_all(hit)) return kWinGroupRevisit;
  }

  // new nodes: prefix chain, suffix chain (:196-199), the aligned pairs'
  for (uint32_t i = t.lane; i < npre; i += t.size) {
    NewNode(g, V0 + i, static_cast<uint8_t>(h.coder[in.seq[i]]));
    g.path[P0 + i] = V0 + i;
  }
  for (uint32_t j = t.lane; j < nsuf; j += t.size) {
    NewNode(g, V0 + npre + j, static_cast<uint8_t>(h.coder[in.seq[last_pos + 1 + j]]));
    g.path[P0 + npre + nvalid + j] = V0 + npre + j;
  }
  for (uint32_t k = t.lane; k < nvalid; k += t.size) {
    const uint32_t v = g.path[P0 + npre + k];
    if (v >= vnew0) NewNode(g, v, static_cast<uint8_t>(h.coder[in.seq[first_pos + k]]));
  }
  t.sync();
  // aligned groups (:228-233): every member lists the new node, the new node
  // lists the members in the order of the aligned-to node's list, then it
  for (uint32_t k = t.lane; k < nfresh; k += t.size) {
    const uint32_t jt = join[k];
    if (jt == kNil) continue;
    const uint32_t n = vnew0 + k;
    const uint32_t cnt = g.al_cnt[jt];
    for (uint32_t i = 0; i < cnt; ++i) {
      const uint32_t kt = g.al[jt * kMaxAligned + i];
      g.al[kt * kMaxAligned + g.al_cnt[kt]] = n;
      g.al_cnt[kt] += 1;
      g.al[n * kMaxAligned + i] = kt;
    }
    g.al[jt * kMaxAligned + cnt] = n;
    g.al_cnt[jt] = static_cast<uint8_t>(cnt + 1);
    g.al[n * kMaxAligned + cnt] = jt;
    g.al_cnt[n] = static_cast<uint8_t>(cnt + 1);
  }
  t.sync();

  // edges in creation order (:81-91, 100-103, 236-244): prefix chain,
  // suffix chain, then one per aligned pair after the first (or after a
  // prefix), then the join of the last aligned node to the suffix.  An
  // edge between two nodes that existed already may exist: it gains weight.
  const uint32_t epre = npre > 0 ? npre - 1 : 0, esuf = nsuf > 0 ? nsuf - 1 : 0;
  for (uint32_t i = t.lane + 1; i < npre; i += t.size) {
    NewEdge(g, E0 + i - 1, V0 + i - 1, V0 + i, static_cast<int64_t>(WeightAt(in, i - 1) + WeightAt(in, i)));
  }
  for (uint32_t j = t.lane + 1; j < nsuf; j += t.size) {
    const uint32_t p = last_pos + 1 + j;
    NewEdge(g, E0 + epre + j - 1, V0 + npre + j - 1, V0 + npre + j,
            static_cast<int64_t>(WeightAt(in, p - 1) + WeightAt(in, p)));
  }
  uint32_t ecount = E0 + epre + esuf;
  for (uint32_t base = 0; base < nvalid; base += t.size) {
    const uint32_t k = base + t.lane;
    bool has = false, fresh = false;
    uint32_t prev = kNil, curr = kNil, found = kNil;
    if (k < nvalid) {
      curr = g.path[P0 + npre + k];
      prev = k > 0 ? g.path[P0 + npre + k - 1] : (npre > 0 ? V0 + npre - 1 : kNil);
      has = prev != kNil;
      if (has && prev < V0 && curr < V0) {
        for (uint32_t e = g.out_head[prev]; e != kNil; e = g.e_next_out[e]) {
          if (e < E0 && g.e_head[e] == curr) {
            found = e;
            break;
          }
        }
      }
      fresh = has && found == kNil;
    }
    uint32_t tot = 0;
    const uint32_t x = t.scan(fresh ? 1u : 0u, &tot);
    // all lanes have read their tail's list before any lane appends to one
    t.sync();
    if (has) {
      const uint32_t p = first_pos + k;
      const int64_t w = static_cast<int64_t>(WeightAt(in, p - 1) + WeightAt(in, p));
      if (fresh) {
        NewEdge(g, ecount + x, prev, curr, w);
      } else {
        g.e_weight[found] += w;
      }
    }
    ecount += tot;
    t.sync();
  }
  if (nsuf > 0 && t.lane == 0) {
    NewEdge(g, ecount, g.path[P0 + npre + nvalid - 1], V0 + npre,
            static_cast<int64_t>(WeightAt(in, last_pos) + WeightAt(in, last_pos + 1)));
  }
  if (nsuf > 0) ecount += 1;
  t.sync();
  if (t.lane == 0) {
    h.V = vnew0 + nfresh;
    h.E = ecount;
    h.path_len = P0 + len;
    g.path_off[s + 1] = P0 + len;
    h.nseq = s + 1;
  }
  t.sync();
  return kWinOk;
}

// Arrays of the topological sort, in LDS (16-bit indices) or in the
// window's global scratch (32-bit).

template <typename Idx>
struct SortSpace {
  Idx* in_off;   // [V + 1] in-edge offsets by node id
  Idx* in_tail;  // [E] in-edge tails, each node's in creation order
  Idx* al_off;   // [V + 1]
  Idx* al_list;  // [A]
  Idx* stack;    // [K]
  uint32_t K;
  Idx* dist;       // [V] source distance by node id (global alignments only)
  Idx* sink;       // [kSortSink]
  uint8_t* marks;  // [V rounded up to 8], 8-byte aligned: 1 visited, 2 finished,
                   // 4 ranked with its group's first member
  uint8_t* sink_mk;  // [kSortSink]
};

// Words (of Idx) of a SortSpace's index arrays.
__host__ __device__ inline uint64_t SortSpaceElems(uint64_t V, uint64_t E, uint64_t A, uint64_t K, bool nw) {
  return (V + 1) + E + (V + 1) + A + K + (nw ? V : 0) + kSortSink;
}

// Bytes of a SortSpace from an 8-byte aligned base: the index arrays, then
// the marks at the next 8-byte boundary, then the marks' sink.
__host__ __device__ inline uint64_t SortSpaceBytes(uint64_t V, uint64_t E, uint64_t A, uint64_t K, bool nw,
                                                   uint64_t idx_bytes) {
  return (SortSpaceElems(V, E, A, K, nw) * idx_bytes + 7) / 8 * 8 + (V + 7) / 8 * 8 + kSortSink;
}

template <typename Idx>
__host__ __device__ inline SortSpace<Idx> CarveSortSpace(void* base, uint32_t V, uint32_t E, uint32_t A,
                                                         uint32_t K, bool nw) {
  SortSpace<Idx> s;
  Idx* p = static_cast<Idx*>(base);
  s.in_off = p;
  p += V + 1;
  s.in_tail = p;
  p += E;
  s.al_off = p;
  p += V + 1;
  s.al_list = p;
  p += A;
  s.stack = p;
  p += K;
  s.K = K;
  s.dist = nw ? p : nullptr;
  p += nw ? V : 0;
  s.sink = p;
  const uint64_t idx_bytes = (SortSpaceElems(V, E, A, K, nw) * sizeof(Idx) + 7) / 8 * 8;
  s.marks = static_cast<uint8_t*>(base) + idx_bytes;
  s.sink_mk = s.marks + (V + 7) / 8 * 8;
  return s;
}

// Bytes 0..7 of w that are zero, as the high bit of each such byte (exact:
// no carry crosses a byte).
__host__ __device__ inline uint64_t ZeroBytes(uint64_t w) {
  constexpr uint64_t k7f = 0x7f7f7f7f7f7f7f7full;
  return ~(((w & k7f) + k7f) | w | k7f);
}

// Graph::TopologicalSort (src/graph.cpp:249-303): depth first over node ids;
// a node is finished once its in-edge tails (and, unless it is ranked with
// another node's group, its aligned nodes) are; a group is ranked when its
// first finished non-ignored member is, the member first, then its aligned
// list.  A node back on top of the stack after its first visit has every
// node that visit pushed finished above it, so it is finished then.  For a
// global alignment each node's source distance (0 without in-edges, else one
// more than its nearest tail's) is set when it finishes, after its tails.
//
// The per-node lists are copied into `sp` first (team-parallel).  The walk
// itself is one serial order, so on a wave its cost is the instructions a
// visit issues (a lone wave issues one every four cycles) plus its dependent
// reads.  Every lane runs it with the same node (its values made uniform, so
// its branches are scalar), and a visit's list work is spread over the
// lanes: lane k takes entry k of the node's in-edge tails followed by its
// aligned nodes (the order the reference pushes them in), reads it and its
// mark, and the unfinished ones go onto the stack in that order through one
// ballot and a prefix count.  Lanes with nothing to write write to a sink
// slot of their own instead of branching.  The rank order is written to the
// window as it forms, its inverse team-parallel at the end.  Returns false
// when the stack of `sp` overflows.
template <class Team, typename Idx>
__host__ __device__ bool TopologicalSort(const Team& t, const WinView& g, const SortSpace<Idx>& sp) {
  const uint32_t V = g.hdr->V, E = g.hdr->E;
  // in-edge lists by node id: offsets, then every edge at its head's offset
  // plus its position in the head's list
  uint32_t run = 0;
  for (uint32_t base = 0; base < V; base += t.size) {
    const uint32_t v = base + t.lane;
    const uint32_t d = v < V ? g.indeg[v] : 0u;
    uint32_t tot = 0;
    const uint32_t x = t.scan(d, &tot);
    if (v < V) sp.in_off[v] = static_cast<Idx>(run + x);
    run += tot;
  }
  if (t.lane == 0) sp.in_off[V] = static_cast<Idx>(run);
  uint32_t arun = 0;
  for (uint32_t base = 0; base < V; base += t.size) {
    const uint32_t v = base + t.lane;
    const uint32_t d = v < V ? g.al_cnt[v] : 0u;
    uint32_t tot = 0;
    const uint32_t x = t.scan(d, &tot);
    if (v < V) {
      sp.al_off[v] = static_cast<Idx>(arun + x);
      sp.marks[v] = 0;
    }
    arun += tot;
  }
  if (t.lane == 0) sp.al_off[V] = static_cast<Idx>(arun);
  t.sync();
  for (uint32_t e = t.lane; e < E; e += t.size) {
    sp.in_tail[static_cast<uint32_t>(sp.in_off[g.e_head[e]]) + g.e_in_pos[e]] = static_cast<Idx>(g.e_tail[e]);
  }
  for (uint32_t v = t.lane; v < V; v += t.size) {
    const uint32_t cnt = g.al_cnt[v], o = sp.al_off[v];
    for (uint32_t k = 0; k < cnt; ++k) sp.al_list[o + k] = static_cast<Idx>(g.al[v * kMaxAligned + k]);
  }
  t.sync();
  const bool nw = sp.dist != nullptr;
  uint8_t* const mk = sp.marks;
  // the rank order goes straight to the window; writes nothing reads go to
  // a sink of 32-bit slots: the window's scratch when the sort space is
  // elsewhere, else the sort space's own sink
  uint32_t* const r2n = g.rank_to_node;
  uint32_t* const rsink = sizeof(Idx) == 4 ? reinterpret_cast<uint32_t*>(sp.sink) : g.scratch;
  auto U = [&](uint32_t x) { return t.uniform(x); };
  const uint32_t Vu = U(V), K = U(sp.K), lane = t.lane;
  uint32_t r = 0, top = 0;
  bool ok = true;
  // entry q of a node's combined list: its in-edge tails, then its aligned nodes
  auto entry = [&](uint32_t q, uint32_t i0, uint32_t nt, uint32_t a0) {
    const Idx* const list = q < nt ? sp.in_tail : sp.al_list;
    return static_cast<uint32_t>(list[q < nt ? i0 + q : a0 + (q - nt)]);
  };
  // the smallest distance over lanes [0, cnt) (cnt > 0)
  auto min_lanes = [&](uint32_t v, uint32_t cnt) { return cnt <= 16 ? t.min16(v) : U(t.min_all(v)); };
  for (uint32_t base = 0; base < Vu && ok; base += 8) {
    uint32_t j = 0;
    uint64_t wu = 0;
    bool word = false;  // wu holds the marks of nodes base..base+7
    while (ok && j < 8) {
      // the next root: the first unmarked node at or after base + j
      if (!word) {
        const uint64_t w = *reinterpret_cast<const uint64_t*>(mk + base);
        wu = static_cast<uint64_t>(U(static_cast<uint32_t>(w >> 32))) << 32 | U(static_cast<uint32_t>(w));
        word = true;
      }
      uint64_t z = ZeroBytes(wu) & (~0ull << (8 * j));
      if (Vu - base < 8) z &= (1ull << (8 * (Vu - base))) - 1;
      if (z == 0) break;
      j = static_cast<uint32_t>(__builtin_ctzll(z)) >> 3;
      uint32_t cur = base + j;
      ++j;
      if (K == 0) {
        ok = false;
        break;
      }
      sp.stack[0] = static_cast<Idx>(cur);
      top = 1;
      bool single = true;  // the root finished at its first visit, pushing nothing
      while (top > 0) {
        const uint32_t m = U(mk[cur]);
        const uint32_t i0 = U(sp.in_off[cur]), i1 = U(sp.in_off[cur + 1]);
        const uint32_t a0 = U(sp.al_off[cur]), a1 = U(sp.al_off[cur + 1]);
        const uint32_t below = U(sp.stack[top >= 2 ? top - 2 : 0]);
        if (m & 2) {  // finished since it was pushed
          --top;
          cur = below;
          continue;
        }
        const bool ign = (m & 4) != 0;
        const uint32_t nt = i1 - i0, n = nt + (ign ? 0u : a1 - a0);
        uint32_t dmin = kNil;
        if ((m & 1) == 0) {
          // first visit: the unfinished tails, then (unless ignored) the
          // unfinished aligned nodes, marked ignored, go onto the stack; the
          // aligned nodes are also written at the ranks they take if the node
          // finishes now (ranks from r on are not final yet)
          bool pushed = false;
          uint32_t last = 0;
          for (uint32_t c = 0; c < n; c += t.size) {
            const uint32_t q = c + lane;
            const bool valid = q < n;
            const uint32_t qq = valid ? q : c;
            const uint32_t val = entry(qq, i0, nt, a0);
            const uint32_t mv = mk[val];
            const bool un = valid && (mv & 2) == 0;
            const uint64_t pm = t.ballot(un);
            if (pm != 0) {
              const uint32_t cnt = static_cast<uint32_t>(__builtin_popcountll(pm));
              if (top + cnt > K) {
                ok = false;
                break;
              }
              *(un ? sp.stack + top + t.below(pm) : sp.sink + lane) = static_cast<Idx>(val);
              *(un && qq >= nt ? mk + val : sp.sink_mk + lane) = static_cast<uint8_t>(mv | 4);
              top += cnt;
              last = t.readlane(val, 63u - static_cast<uint32_t>(__builtin_clzll(pm)));
              pushed = true;
            }
            if (!ign) *(valid && qq >= nt ? r2n + r + 1 + (qq - nt) : rsink + lane) = val;
            if (nw && c < nt) {
              const uint32_t dv = sp.dist[val];
              const uint32_t dm = min_lanes(valid && qq < nt ? dv : kNil, nt - c);
              dmin = dm < dmin ? dm : dmin;
            }
          }
          if (!ok) break;
          if (pushed) {
            mk[cur] = static_cast<uint8_t>(m | 1);
            cur = last;
            single = false;
            continue;
          }
        } else {
          // back on top, every node it pushed finished: its tails' distances
          // and the aligned nodes it ranks
          for (uint32_t c = nw ? 0u : nt; c < n; c += t.size) {
            const uint32_t q = c + lane;
            const bool valid = q < n;
            const uint32_t qq = valid ? q : c;
            const uint32_t val = entry(qq, i0, nt, a0);
            if (!ign) *(valid && qq >= nt ? r2n + r + 1 + (qq - nt) : rsink + lane) = val;
            if (nw && c < nt) {
              const uint32_t dv = sp.dist[val];
              const uint32_t dm = min_lanes(valid && qq < nt ? dv : kNil, nt - c);
              dmin = dm < dmin ? dm : dmin;
            }
          }
        }
        mk[cur] = static_cast<uint8_t>(m | 2);
        if (nw) sp.dist[cur] = static_cast<Idx>(nt > 0 ? dmin + 1 : 0u);
        if (!ign) {
          r2n[r] = cur;
          r += 1 + (a1 - a0);
        }
        --top;
        cur = below;
      }
      // a root that finished alone changed no mark but its own
      if (single) {
        wu |= 2ull << (8 * (j - 1));
      } else {
        word = false;
      }
    }
  }
  t.sync();
  // the inverse of the rank order
  for (uint32_t q = t.lane; q < V; q += t.size) g.rank_of[r2n[q]] = q;
  t.sync();
  return t.any(!ok) ? false : true;
}

// The graph block of the alignment kernels (MakeView layout) from the rank
// order: per rank its row record (base | sink << 8 | profile slot << 9 |
// predecessor count << 12, first predecessor row, offset of its predecessor
// list, first-column H), node ids, predecessor offsets, first columns
// (reference simd...:568-577, 590-599, 607-694) and predecessor rows (rank +
// 1, in-edge order).  First columns of a global alignment follow the source
// distance d of each node (the fewest edges from a source, which
// TopologicalSort leaves in sp.dist): the vertical gap column is g + d * e
// (q + d * c), the linear H column (d + 1) * g.
// Also sets the header's R, nslots and slot_chars.
template <class Team, typename Idx>
__host__ __device__ void BuildBlock(const Team& t, const WinView& g, const SortSpace<Idx>& sp, uint32_t* block,
                                    int32_t type, int32_t subtype, int32_t pg, int32_t pe, int32_t pq,
                                    int32_t pc) {
  WinHeader& h = *g.hdr;
  const uint32_t V = h.V;
  uint32_t* rows = block;
  int32_t* node_id = reinterpret_cast<int32_t*>(block + 4ull * V);
  uint32_t* poff = block + 5ull * V;
  int32_t* fcH = reinterpret_cast<int32_t*>(block + 6ull * V + 1);
  int32_t* fcF = fcH + (V + 1);
  int32_t* fcO = fcF + (V + 1);
  uint32_t* preds = block + 9ull * V + 4;
  const bool nw = type == 1;
  const uint32_t nslots = h.ncodes <= 4 ? h.ncodes : 0u;
  uint32_t run = 0, rsum = 0;
  for (uint32_t base = 0; base < V; base += t.size) {
    const uint32_t r = base + t.lane;
    const bool live = r < V;
    const uint32_t v = live ? g.rank_to_node[r] : 0u;
    const uint32_t np = live ? g.indeg[v] : 0u;
    uint32_t tot = 0;
    const uint32_t x = t.scan(np, &tot);
    if (live) {
      const uint32_t pb = run + x;
      const uint32_t i0 = sp.in_off[v];
      for (uint32_t i = 0; i < np; ++i) preds[pb + i] = g.rank_of[sp.in_tail[i0 + i]] + 1;
      int32_t fh = 0;
      if (nw) {
        const int32_t d = static_cast<int32_t>(static_cast<uint32_t>(sp.dist[v]));
        const int32_t ff = pg + d * pe, fo = pq + d * pc;
        fcF[r + 1] = ff;
        fcO[r + 1] = fo;
        fh = subtype == 2 ? (ff > fo ? ff : fo) : (subtype == 1 ? ff : (d + 1) * pg);
      } else {
        fcF[r + 1] = 0;
        fcO[r + 1] = 0;
      }
      fcH[r + 1] = fh;
      const uint32_t slot = nslots ? static_cast<uint32_t>(g.code[v]) : 7u;
      const uint32_t n1 = np > 0 ? np : 1u;
      rows[4ull * r + 0] = static_cast<uint32_t>(h.decoder[g.code[v]]) | (g.outdeg[v] == 0 ? 0x100u : 0u) |
                           (slot << 9) | (n1 << 12);
      rows[4ull * r + 1] = np > 0 ? g.rank_of[sp.in_tail[i0]] + 1 : 0u;
      rows[4ull * r + 2] = pb;
      rows[4ull * r + 3] = static_cast<uint32_t>(fh);
      node_id[r] = static_cast<int32_t>(v);
      poff[r] = pb;
      rsum += n1;
    }
    run += tot;
  }
  const uint32_t R = t.sum_all(rsum);
  if (t.lane == 0) {
    poff[V] = run;
    fcH[0] = 0;
    fcF[0] = 0;
    fcO[0] = 0;
    h.R = R;
    h.nslots = nslots;
    uint32_t chars = 0;
    for (uint32_t c = 0; c < nslots; ++c) chars |= static_cast<uint32_t>(h.decoder[c]) << (8 * c);
    h.slot_chars = chars;
  }
  t.sync();
}

// ------------------------------------------------ consensus and MSA ---

// Graph::TraverseHeaviestBundle + BranchCompletion (src/graph.cpp:433-516):
// per node in rank order the heaviest in-edge (ties to a later edge whose
// tail scores at least as much), the score carried from it, the first best
// node; while that node has out-edges the walk restarts after it with every
// other tail into its successors cut off.  Writes the consensus node ids
// into out (in order) and returns their count.  The in-edges by node id
// come first (team), the walk is serial (lane 0).  Scratch: 4V + E + 1
// words of global memory.
template <class Team>
__host__ __device__ uint32_t Consensus(const Team& t, const WinView& g, uint32_t* scratch, uint32_t* out) {
  const uint32_t V = g.hdr->V, E = g.hdr->E;
  if (V == 0) return 0;
  uint32_t* in_off = scratch;        // [V + 1]
  uint32_t* in_edge = in_off + V + 1;  // [E]
  uint32_t* pred = in_edge + E;      // [V]
  int64_t* score = reinterpret_cast<int64_t*>(pred + V + ((V + E + 1 + V) & 1u));  // [V], 8-byte aligned
  uint32_t run = 0;
  for (uint32_t base = 0; base < V; base += t.size) {
    const uint32_t v = base + t.lane;
    const uint32_t d = v < V ? g.indeg[v] : 0u;
    uint32_t tot = 0;
    const uint32_t x = t.scan(d, &tot);
    if (v < V) in_off[v] = run + x;
    run += tot;
  }
  if (t.lane == 0) in_off[V] = run;
  t.sync();
  for (uint32_t e = t.lane; e < E; e += t.size) in_edge[in_off[g.e_head[e]] + g.e_in_pos[e]] = e;
  for (uint32_t v = t.lane; v < V; v += t.size) {
    pred[v] = kNil;
    score[v] = -1;
  }
  t.sync();
  uint32_t n = 0;
  if (t.lane == 0) {
    auto relax = [&](uint32_t v, bool skip_cut) {
      for (uint32_t i = in_off[v]; i < in_off[v + 1]; ++i) {
        const uint32_t e = in_edge[i], tl = g.e_tail[e];
        if (skip_cut && score[tl] == -1) continue;
        const int64_t w = g.e_weight[e];
        if (score[v] < w || (score[v] == w && score[pred[v]] <= score[tl])) {
          score[v] = w;
          pred[v] = tl;
        }
      }
      if (pred[v] != kNil) score[v] += score[pred[v]];
    };
    uint32_t best = kNil;
    for (uint32_t r = 0; r < V; ++r) {
      const uint32_t v = g.rank_to_node[r];
      relax(v, false);
      if (best == kNil || score[best] < score[v]) best = v;
    }
    while (g.outdeg[best] != 0) {
      const uint32_t start = best, rank = g.rank_of[start];
      for (uint32_t e = g.out_head[start]; e != kNil; e = g.e_next_out[e]) {
        const uint32_t hd = g.e_head[e];
        for (uint32_t i = in_off[hd]; i < in_off[hd + 1]; ++i) {
          const uint32_t tl = g.e_tail[in_edge[i]];
          if (tl != start) score[tl] = -1;
        }
      }
      best = kNil;
      for (uint32_t r = rank + 1; r < V; ++r) {
        const uint32_t v = g.rank_to_node[r];
        score[v] = -1;
        pred[v] = kNil;
        relax(v, true);
        if (best == kNil || score[best] < score[v]) best = v;
      }
    }
    for (uint32_t v = best; v != kNil; v = pred[v]) out[n++] = v;
    for (uint32_t i = 0; i < n / 2; ++i) {
      const uint32_t x = out[i];
      out[i] = out[n - 1 - i];
      out[n - 1 - i] = x;
    }
  }
  n = t.max_all(n);
  t.sync();
  return n;
}

// Graph::InitializeMultipleSequenceAlignment (src/graph.cpp:321-337): one
// column per aligned group, in rank order; a group's first-ranked member
// opens its column.  Writes col[node id]; returns the width.
template <class Team>
__host__ __device__ uint32_t MsaColumns(const Team& t, const WinView& g, uint32_t* col) {
  const uint32_t V = g.hdr->V;
  uint32_t run = 0;
  for (uint32_t base = 0; base < V; base += t.size) {
    const uint32_t r = base + t.lane;
    uint32_t lead = 0, v = 0;
    if (r < V) {
      v = g.rank_to_node[r];
      lead = 1;
      for (uint32_t k = 0; k < g.al_cnt[v]; ++k) {
        if (g.rank_of[g.al[v * kMaxAligned + k]] < r) lead = 0;
      }
    }
    uint32_t tot = 0;
    const uint32_t x = t.scan(lead, &tot);
    if (r < V) col[v] = run + x + lead - 1;
    run += tot;
  }
  t.sync();
  return run;
}

// Graph::GenerateMultipleSequenceAlignment (src/graph.cpp:339-366): row s
// holds sequence s's bases at its nodes' columns, '-' elsewhere; with
// cons_len > 0 a last row holds the consensus.  rows: (nseq [+ 1]) * width.
template <class Team>
__host__ __device__ void WriteMsa(const Team& t, const WinView& g, const uint32_t* col, uint32_t width,
                                  const uint32_t* cons, uint32_t cons_len, uint8_t* rows) {
  const uint32_t S = g.hdr->nseq + (cons_len > 0 ? 1u : 0u);
  const uint64_t cells = static_cast<uint64_t>(S) * width;
  for (uint64_t i = t.lane; i < cells; i += t.size) rows[i] = '-';
  t.sync();
  const uint32_t P = g.hdr->path_len;
  for (uint32_t s = 0; s < g.hdr->nseq; ++s) {
    const uint32_t a = g.path_off[s], b = g.path_off[s + 1];
    for (uint32_t k = a + t.lane; k < b; k += t.size) {
      const uint32_t v = g.path[k];
      rows[static_cast<uint64_t>(s) * width + col[v]] = g.hdr->decoder[g.code[v]];
    }
  }
  for (uint32_t k = t.lane; k < cons_len; k += t.size) {
    const uint32_t v = cons[k];
    rows[static_cast<uint64_t>(g.hdr->nseq) * width + col[v]] = g.hdr->decoder[g.code[v]];
  }
  (void)P;
  t.sync();
}

}  // namespace wg
}  // namespace spoa_amd
