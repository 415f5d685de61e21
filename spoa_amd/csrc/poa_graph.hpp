// A flat partial-order graph with the reference Graph's semantics, for the
// batched window driver.
//
// spoa::Graph (include/spoa/graph.hpp) keeps the reference's public shape:
// heap nodes and edges linked by pointers.  The batched form merges
// thousands of windows per round and only needs what the alignment and
// the outputs read, so this graph stores the same information in flat
// arrays: node codes, edges with per-head and per-tail lists threaded in
// creation order (the order of Node::inedges / Node::outedges), aligned
// groups, and each sequence's path instead of per-edge labels (a read
// visits a node at most once, so following its label from its first node,
// src/graph.cpp:22-30 and 339-366, walks exactly its path).
//
// Every operation restates the reference function it replaces (cited per
// function) and is checked against the oracle in tests/test_host_graph.py.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "spoa_hip.h"

namespace spoa_amd {

class PoaGraph {
 public:
  PoaGraph();

  void Clear();
  std::uint32_t num_nodes() const { return static_cast<std::uint32_t>(code_.size()); }
  std::uint32_t num_codes() const { return num_codes_; }

  // Graph::AddAlignment (src/graph.cpp:155-247): pairs (node id or -1,
  // position or -1) in path order; weights[i] per base.
  void AddAlignment(const std::int32_t* pairs, std::uint32_t n_pairs, const char* seq, std::uint32_t len,
                    const std::uint32_t* weights);

  // Graph::GenerateConsensus (src/graph.cpp:368-375)
  std::string GenerateConsensus();
  // Graph::GenerateMultipleSequenceAlignment (src/graph.cpp:339-366)
  std::vector<std::string> GenerateMultipleSequenceAlignment(bool include_consensus = false);

  const std::vector<std::uint32_t>& rank_to_node() const { return rank_to_node_; }

  // The rank-ordered view the alignment kernels take (refreshed by every
  // topological sort).
  spoa_hip_graph View() const;

 private:
  static constexpr std::uint32_t kNil = 0xffffffffu;

  std::uint32_t AddNode(std::uint32_t code);
  void AddEdge(std::uint32_t tail, std::uint32_t head, std::uint32_t weight);
  std::int64_t AddSequence(const char* seq, const std::uint32_t* w, std::uint32_t begin, std::uint32_t end,
                           std::vector<std::uint32_t>* path);
  void TopologicalSort();
  void RebuildView();
  void TraverseHeaviestBundle();
  std::uint32_t BranchCompletion(std::uint32_t rank, std::vector<std::int64_t>& score,
                                 std::vector<std::uint32_t>& pred);
  std::vector<std::uint32_t> MsaColumns(std::uint32_t* width) const;

  std::int32_t coder_[256];
  std::int32_t decoder_[256];
  std::uint32_t num_codes_ = 0;

  // nodes
  std::vector<std::uint8_t> code_;
  std::vector<std::uint32_t> first_in_, last_in_, first_out_, last_out_;
  std::vector<std::vector<std::uint32_t>> aligned_;
  // edges, threaded per head / per tail in creation order
  std::vector<std::uint32_t> tail_, head_, next_in_, next_out_;
  std::vector<std::int64_t> weight_;
  // sequences: the nodes each one visits, in order
  std::vector<std::vector<std::uint32_t>> paths_;

  std::vector<std::uint32_t> rank_to_node_;
  std::vector<std::uint32_t> consensus_;

  // kernel view (rank order)
  std::vector<std::uint8_t> v_char_, v_sink_;
  std::vector<std::int32_t> v_id_;
  std::vector<std::uint32_t> v_off_, v_pred_, rank_of_;

  // scratch of the topological sort
  std::vector<std::uint8_t> marks_, ignored_;
  std::vector<std::uint32_t> stack_;
};

}  // namespace spoa_amd
