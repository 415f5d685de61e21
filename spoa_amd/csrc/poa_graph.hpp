// A flat partial-order graph with the reference Graph's semantics, for the
// batched window driver.
//
// spoa::Graph (include/spoa/graph.hpp) keeps the reference's public shape:
// heap nodes and edges linked by pointers.  The batched form merges
// thousands of windows per round and only needs what the alignment and
// the outputs read, so this graph stores the same information in a fixed
// set of flat arrays: node codes; edges; per node its in-edges, out-edges
// and aligned nodes as lists in creation order (the order of
// Node::inedges / Node::outedges / Node::aligned_nodes), each list a block
// of one shared pool; and each sequence's path instead of per-edge labels
// (a read visits a node at most once, so following its label from its
// first node, src/graph.cpp:22-30 and 339-366, walks exactly its path).
// Clear() keeps every array's capacity, so a graph reused for the next
// window allocates nothing once it has seen a window of that size.
//
// Every operation restates the reference function it replaces (cited per
// function) and is checked against the oracle in tests/test_host_graph.py.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "spoa_hip.h"

namespace spoa_amd {

// Per-owner lists of T in one pool: a list is a block [start, start +
// size) of capacity cap; a full list moves to a block of twice the capacity
// at the end of the pool (or grows in place when it is the last block).
template <typename T>
class BlockLists {
 public:
  void clear() {
    pool_.clear();
    meta_.clear();
  }
  void add_owner() { meta_.push_back(Meta{0, 0, 0}); }
  std::uint32_t size(std::uint32_t o) const { return meta_[o].size; }
  const T* begin(std::uint32_t o) const { return pool_.data() + meta_[o].start; }
  const T* end(std::uint32_t o) const { return pool_.data() + meta_[o].start + meta_[o].size; }
  T at(std::uint32_t o, std::uint32_t i) const { return pool_[meta_[o].start + i]; }
  void push(std::uint32_t o, T v) {
    Meta& m = meta_[o];
    if (m.size == m.cap) Grow(m);
    pool_[m.start + m.size++] = v;
  }

 private:
  struct Meta {
    std::uint32_t start, size, cap;
  };
  void Grow(Meta& m) {
    const std::uint32_t cap = m.cap ? 2 * m.cap : 2;
    const std::size_t end = pool_.size();
    if (m.cap != 0 && m.start + m.cap == end) {
      pool_.resize(end + (cap - m.cap));
    } else {
      pool_.resize(end + cap);
      for (std::uint32_t i = 0; i < m.size; ++i) pool_[end + i] = pool_[m.start + i];
      m.start = static_cast<std::uint32_t>(end);
    }
    m.cap = cap;
  }
  std::vector<T> pool_;
  std::vector<Meta> meta_;
};

// One entry of a node's in-edge (out-edge) list: the edge's other end --
// its tail (head) -- kept beside the edge id, so walks over the lists read
// the neighbour without a lookup by edge.
struct Adj {
  std::uint32_t node, edge;
};

class PoaGraph {
 public:
  PoaGraph();

  void Clear();
  std::uint32_t num_nodes() const { return static_cast<std::uint32_t>(code_.size()); }
  std::uint32_t num_codes() const { return num_codes_; }
  std::uint32_t num_sequences() const { return static_cast<std::uint32_t>(path_off_.size() - 1); }

  // Graph::AddAlignment (src/graph.cpp:155-247): pairs (node id or -1,
  // position or -1) in path order; weights[i] per base.
  void AddAlignment(const std::int32_t* pairs, std::uint32_t n_pairs, const char* seq, std::uint32_t len,
                    const std::uint32_t* weights);

  // Graph::GenerateConsensus (src/graph.cpp:368-375)
  std::string GenerateConsensus();
  // Graph::GenerateMultipleSequenceAlignment (src/graph.cpp:339-366)
  std::vector<std::string> GenerateMultipleSequenceAlignment(bool include_consensus = false);
  // the same rows one after another in one block: n_rows rows of width
  // characters
  void MultipleSequenceAlignmentBlock(bool include_consensus, std::string* block, std::uint32_t* n_rows,
                                      std::uint32_t* width);

  const std::vector<std::uint32_t>& rank_to_node() const { return rank_to_node_; }

  // The rank-ordered view the alignment kernels take (refreshed by every
  // topological sort).
  spoa_hip_graph View() const;

 private:
  static constexpr std::uint32_t kNil = 0xffffffffu;

  std::uint32_t AddNode(std::uint32_t code);
  void AddEdge(std::uint32_t tail, std::uint32_t head, std::uint32_t weight);
  std::int64_t AddSequence(const char* seq, const std::uint32_t* w, std::uint32_t begin, std::uint32_t end,
                           bool on_path);
  void TopologicalSort();
  void TraverseHeaviestBundle();
  std::uint32_t BranchCompletion(std::uint32_t rank, std::vector<std::int64_t>& score,
                                 std::vector<std::uint32_t>& pred);
  std::vector<std::uint32_t> MsaColumns(std::uint32_t* width) const;

  std::int32_t coder_[256];
  std::int32_t decoder_[256];
  std::uint32_t num_codes_ = 0;

  // nodes: code, in-edges, out-edges, aligned nodes (node ids)
  std::vector<std::uint8_t> code_;
  BlockLists<Adj> in_, out_;
  BlockLists<std::uint32_t> aligned_;
  // edges (their ends live in the in- and out-edge lists)
  std::vector<std::int64_t> weight_;
  // sequences: sequence s visits path_[path_off_[s] .. path_off_[s + 1])
  std::vector<std::uint32_t> path_, path_off_;

  std::vector<std::uint32_t> rank_to_node_;
  std::vector<std::uint32_t> consensus_;

  // kernel view (rank order)
  std::vector<std::uint8_t> v_char_, v_sink_;
  std::vector<std::int32_t> v_id_;
  std::vector<std::uint32_t> v_off_, v_pred_, rank_of_;

  // scratch of the topological sort
  std::vector<std::uint8_t> marks_;
  std::vector<std::uint32_t> stack_;
};

}  // namespace spoa_amd
