// spoa::Graph -- drop-in for the reference POA graph
// (include/spoa/graph.hpp:26-318 of spoa v4.0.8): the same public node and
// edge records, accessors, AddAlignment overloads, consensus and MSA.
// Also the windowed-alignment helpers (Subgraph, UpdateAlignment), the
// Graphviz writer (PrintDot) and the coverage-summary GenerateConsensus.
//
// Beyond the reference, the graph keeps a mutation counter and a
// rank-ordered CSR view (predecessor ranks in in-edge order, decoded bases,
// sink flags) refreshed once per topological sort; the GPU engine hands
// that view to the device as is.
#ifndef SPOA_AMD_GRAPH_HPP_
#define SPOA_AMD_GRAPH_HPP_

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace spoa {

using Alignment = std::vector<std::pair<std::int32_t, std::int32_t>>;

class Graph {
 public:
  Graph();

  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;

  Graph(Graph&&) = default;
  Graph& operator=(Graph&&) = default;

  ~Graph() = default;

  struct Node;
  struct Edge;

  struct Node {
    Node(std::uint32_t id, std::uint32_t code);

    Node* Successor(std::uint32_t label) const;
    std::uint32_t Coverage() const;

    std::uint32_t id;
    std::uint32_t code;
    std::vector<Edge*> inedges;
    std::vector<Edge*> outedges;
    std::vector<Node*> aligned_nodes;
  };

  struct Edge {
    Edge(Node* tail, Node* head, std::uint32_t label, std::uint32_t weight);

    void AddSequence(std::uint32_t label, std::uint32_t weight = 1);

    Node* tail;
    Node* head;
    std::vector<std::uint32_t> labels;
    std::int64_t weight;
  };

  const std::vector<std::unique_ptr<Node>>& nodes() const { return nodes_; }
  const std::vector<std::unique_ptr<Edge>>& edges() const { return edges_; }
  const std::vector<Node*>& rank_to_node() const { return rank_to_node_; }
  const std::vector<Node*>& sequences() const { return sequences_; }
  std::uint32_t num_codes() const { return num_codes_; }
  std::uint8_t coder(std::uint8_t c) const { return coder_[c]; }
  std::uint8_t decoder(std::uint8_t code) const { return decoder_[code]; }
  const std::vector<Node*>& consensus() const { return consensus_; }

  void AddAlignment(const Alignment& alignment, const std::string& sequence,
                    std::uint32_t weight = 1);
  void AddAlignment(const Alignment& alignment, const std::string& sequence,
                    const std::vector<std::uint32_t>& weights);
  void AddAlignment(const Alignment& alignment, const std::string& sequence,
                    const std::string& quality);
  void AddAlignment(const Alignment& alignment, const char* sequence,
                    std::uint32_t sequence_len, std::uint32_t weight = 1);
  void AddAlignment(const Alignment& alignment, const char* sequence,
                    std::uint32_t sequence_len,
                    const std::vector<std::uint32_t>& weights);
  void AddAlignment(const Alignment& alignment, const char* sequence,
                    std::uint32_t sequence_len, const char* quality,
                    std::uint32_t quality_len);

  std::vector<std::string> GenerateMultipleSequenceAlignment(
      bool include_consensus = false);

  std::string GenerateConsensus();

  std::string GenerateConsensus(std::vector<std::uint32_t>* summary, bool verbose = false);

  Graph Subgraph(std::uint32_t begin, std::uint32_t end, std::vector<const Node*>* subgraph_to_graph) const;

  void UpdateAlignment(const std::vector<const Node*>& subgraph_to_graph, Alignment* alignment) const;

  // print with Graphviz
  void PrintDot(const std::string& path) const;

  void Clear();

  // Rank-ordered CSR view used by the GPU engine (valid until the next
  // mutation).  pred_rank lists predecessor ranks in in-edge order.
  struct RankView {
    std::vector<std::uint8_t> row_char;
    std::vector<std::int32_t> node_id;
    std::vector<std::uint32_t> pred_off;
    std::vector<std::uint32_t> pred_rank;
    std::vector<std::uint8_t> is_sink;
  };
  const RankView& rank_view() const { return view_; }
  std::uint64_t version() const { return version_; }

 private:
  Node* AddNode(std::uint32_t code);
  void AddEdge(Node* tail, Node* head, std::uint32_t weight);
  Node* AddSequence(const char* sequence,
                    const std::vector<std::uint32_t>& weights,
                    std::uint32_t begin, std::uint32_t end);
  void TopologicalSort();
  void RebuildRankView();
  void TraverseHeaviestBundle();
  Node* BranchCompletion(std::uint32_t rank, std::vector<std::int64_t>* scores,
                         std::vector<Node*>* predecessors);
  std::vector<std::uint32_t> InitializeMultipleSequenceAlignment(
      std::uint32_t* row_size = nullptr) const;
  std::vector<bool> ExtractSubgraph(const Node* from, const Node* floor) const;

  std::uint32_t num_codes_;
  std::vector<std::int32_t> coder_;
  std::vector<std::int32_t> decoder_;
  std::vector<Node*> sequences_;
  std::vector<std::unique_ptr<Node>> nodes_;
  std::vector<std::unique_ptr<Edge>> edges_;
  std::vector<Node*> rank_to_node_;
  std::vector<Node*> consensus_;
  RankView view_;
  std::uint64_t version_ = 0;
};

}  // namespace spoa

#endif  // SPOA_AMD_GRAPH_HPP_
