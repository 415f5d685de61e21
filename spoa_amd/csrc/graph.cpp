// spoa::Graph for the MI355X alignment path.  Behaviour follows the
// reference src/graph.cpp (spoa v4.0.8) function by function; each
// function cites the lines it reproduces.  After every topological sort the
// graph also refreshes a rank-ordered CSR view that the GPU engine uploads
// without further host work.

#include "spoa/graph.hpp"

#include <algorithm>
#include <fstream>
#include <stdexcept>
#include <unordered_set>

namespace spoa {

Graph::Node::Node(std::uint32_t id_, std::uint32_t code_) : id(id_), code(code_) {}

// src/graph.cpp:22-30: first out-edge carrying the label
Graph::Node* Graph::Node::Successor(std::uint32_t label) const {
  for (Edge* e : outedges) {
    if (std::find(e->labels.begin(), e->labels.end(), label) != e->labels.end()) return e->head;
  }
  return nullptr;
}

// src/graph.cpp:32-47
std::uint32_t Graph::Node::Coverage() const {
  std::unordered_set<std::uint32_t> seen;
  for (Edge* e : inedges) seen.insert(e->labels.begin(), e->labels.end());
  for (Edge* e : outedges) seen.insert(e->labels.begin(), e->labels.end());
  return static_cast<std::uint32_t>(seen.size());
}

Graph::Edge::Edge(Node* t, Node* h, std::uint32_t label, std::uint32_t w)
    : tail(t), head(h), labels(1, label), weight(w) {}

// src/graph.cpp:60-63
void Graph::Edge::AddSequence(std::uint32_t label, std::uint32_t w) {
  labels.push_back(label);
  weight += w;
}

Graph::Graph() : num_codes_(0), coder_(256, -1), decoder_(256, -1) {}

Graph::Node* Graph::AddNode(std::uint32_t code) {
  nodes_.emplace_back(new Node(static_cast<std::uint32_t>(nodes_.size()), code));
  return nodes_.back().get();
}

// src/graph.cpp:81-91: parallel edges merge; the label is the index of the
// sequence being added.
void Graph::AddEdge(Node* tail, Node* head, std::uint32_t weight) {
  const std::uint32_t label = static_cast<std::uint32_t>(sequences_.size());
  for (Edge* e : tail->outedges) {
    if (e->head == head) {
      e->AddSequence(label, weight);
      return;
    }
  }
  edges_.emplace_back(new Edge(tail, head, label, weight));
  tail->outedges.push_back(edges_.back().get());
  head->inedges.push_back(edges_.back().get());
}

// src/graph.cpp:93-110: a chain of fresh nodes for sequence[begin, end)
Graph::Node* Graph::AddSequence(const char* sequence, const std::vector<std::uint32_t>& weights,
                                std::uint32_t begin, std::uint32_t end) {
  if (begin == end) return nullptr;
  Node* first = nullptr;
  Node* prev = nullptr;
  for (std::uint32_t i = begin; i < end; ++i) {
    Node* curr = AddNode(coder_[static_cast<std::uint8_t>(sequence[i])]);
    if (prev) AddEdge(prev, curr, weights[i - 1] + weights[i]);
    if (!first) first = curr;
    prev = curr;
  }
  return first;
}

void Graph::AddAlignment(const Alignment& alignment, const std::string& sequence, std::uint32_t weight) {
  AddAlignment(alignment, sequence.c_str(), static_cast<std::uint32_t>(sequence.size()), weight);
}

void Graph::AddAlignment(const Alignment& alignment, const char* sequence, std::uint32_t sequence_len,
                         std::uint32_t weight) {
  std::vector<std::uint32_t> weights(sequence_len, weight);
  AddAlignment(alignment, sequence, sequence_len, weights);
}

void Graph::AddAlignment(const Alignment& alignment, const std::string& sequence, const std::string& quality) {
  AddAlignment(alignment, sequence.c_str(), static_cast<std::uint32_t>(sequence.size()), quality.c_str(),
               static_cast<std::uint32_t>(quality.size()));
}

// src/graph.cpp:133-145: Phred+33 qualities become weights
void Graph::AddAlignment(const Alignment& alignment, const char* sequence, std::uint32_t sequence_len,
                         const char* quality, std::uint32_t quality_len) {
  std::vector<std::uint32_t> weights(quality_len);
  for (std::uint32_t i = 0; i < quality_len; ++i) weights[i] = static_cast<std::uint32_t>(quality[i] - 33);
  AddAlignment(alignment, sequence, sequence_len, weights);
}

void Graph::AddAlignment(const Alignment& alignment, const std::string& sequence,
                         const std::vector<std::uint32_t>& weights) {
  AddAlignment(alignment, sequence.c_str(), static_cast<std::uint32_t>(sequence.size()), weights);
}

// src/graph.cpp:155-247
void Graph::AddAlignment(const Alignment& alignment, const char* sequence, std::uint32_t sequence_len,
                         const std::vector<std::uint32_t>& weights) {
  if (sequence_len == 0) return;
  if (sequence_len != weights.size()) {
    throw std::invalid_argument(
        "[spoa::Graph::AddAlignment] error: "
        "sequence and weights are of unequal size!");
  }
  for (std::uint32_t i = 0; i < sequence_len; ++i) {
    const std::uint8_t ch = static_cast<std::uint8_t>(sequence[i]);
    if (coder_[ch] == -1) {
      coder_[ch] = static_cast<std::int32_t>(num_codes_);
      decoder_[num_codes_++] = sequence[i];
    }
  }
  ++version_;

  if (alignment.empty()) {
    sequences_.push_back(AddSequence(sequence, weights, 0, sequence_len));
    TopologicalSort();
    return;
  }

  // positions of the aligned bases, validated
  std::int32_t first_pos = -1, last_pos = -1;
  for (const auto& pr : alignment) {
    if (pr.second == -1) continue;
    if (pr.second < 0 || pr.second >= static_cast<std::int32_t>(sequence_len)) {
      throw std::invalid_argument("[spoa::Graph::AddAlignment] error: invalid alignment");
    }
    if (first_pos == -1) first_pos = pr.second;
    last_pos = pr.second;
  }
  if (first_pos == -1) {
    throw std::invalid_argument("[spoa::Graph::AddAlignment] error: missing sequence in alignment");
  }

  // unaligned prefix and suffix become fresh chains
  Node* begin = AddSequence(sequence, weights, 0, static_cast<std::uint32_t>(first_pos));
  Node* prev = begin ? nodes_.back().get() : nullptr;
  Node* last = AddSequence(sequence, weights, static_cast<std::uint32_t>(last_pos) + 1, sequence_len);

  for (const auto& pr : alignment) {
    if (pr.second == -1) continue;
    const std::uint32_t code = static_cast<std::uint32_t>(coder_[static_cast<std::uint8_t>(sequence[pr.second])]);
    Node* curr = nullptr;
    if (pr.first == -1) {
      curr = AddNode(code);
    } else {
      Node* anchor = nodes_[pr.first].get();
      if (anchor->code == code) {
        curr = anchor;
      } else {
        for (Node* a : anchor->aligned_nodes) {
          if (a->code == code) {
            curr = a;
            break;
          }
        }
        if (!curr) {
          // join the anchor's aligned group (every member links to every other)
          curr = AddNode(code);
          for (Node* a : anchor->aligned_nodes) {
            a->aligned_nodes.push_back(curr);
            curr->aligned_nodes.push_back(a);
          }
          anchor->aligned_nodes.push_back(curr);
          curr->aligned_nodes.push_back(anchor);
        }
      }
    }
    if (!begin) begin = curr;
    if (prev) AddEdge(prev, curr, weights[pr.second - 1] + weights[pr.second]);
    prev = curr;
  }
  if (last) AddEdge(prev, last, weights[last_pos] + weights[last_pos + 1]);
  sequences_.push_back(begin);
  TopologicalSort();
}

// src/graph.cpp:249-303: depth-first post-order over node ids; a node is
// finished once all its in-edge tails (and, for the group's representative,
// all aligned nodes) are finished; the representative is ranked together
// with its whole aligned group.
void Graph::TopologicalSort() {
  const std::size_t V = nodes_.size();
  rank_to_node_.clear();
  rank_to_node_.reserve(V);
  std::vector<std::uint8_t> state(V, 0);  // 0 new, 1 open, 2 done
  std::vector<std::uint8_t> deferred(V, 0);
  std::vector<Node*> stack;
  for (const auto& root : nodes_) {
    if (state[root->id] != 0) continue;
    stack.push_back(root.get());
    while (!stack.empty()) {
      Node* curr = stack.back();
      bool ready = true;
      if (state[curr->id] != 2) {
        for (Edge* e : curr->inedges) {
          if (state[e->tail->id] != 2) {
            stack.push_back(e->tail);
            ready = false;
          }
        }
        if (!deferred[curr->id]) {
          for (Node* a : curr->aligned_nodes) {
            if (state[a->id] != 2) {
              stack.push_back(a);
              deferred[a->id] = 1;
              ready = false;
            }
          }
        }
        if (ready) {
          state[curr->id] = 2;
          if (!deferred[curr->id]) {
            rank_to_node_.push_back(curr);
            for (Node* a : curr->aligned_nodes) rank_to_node_.push_back(a);
          }
        } else {
          state[curr->id] = 1;
        }
      }
      if (ready) stack.pop_back();
    }
  }
  RebuildRankView();
}

void Graph::RebuildRankView() {
  const std::size_t V = rank_to_node_.size();
  std::vector<std::uint32_t> rank_of(nodes_.size(), 0);
  for (std::size_t r = 0; r < V; ++r) rank_of[rank_to_node_[r]->id] = static_cast<std::uint32_t>(r);
  view_.row_char.resize(V);
  view_.node_id.resize(V);
  view_.is_sink.resize(V);
  view_.pred_off.resize(V + 1);
  view_.pred_rank.clear();
  view_.pred_rank.reserve(edges_.size());
  for (std::size_t r = 0; r < V; ++r) {
    const Node* nd = rank_to_node_[r];
    view_.pred_off[r] = static_cast<std::uint32_t>(view_.pred_rank.size());
    for (const Edge* e : nd->inedges) view_.pred_rank.push_back(rank_of[e->tail->id]);
    view_.row_char[r] = static_cast<std::uint8_t>(decoder_[nd->code]);
    view_.node_id[r] = static_cast<std::int32_t>(nd->id);
    view_.is_sink[r] = nd->outedges.empty() ? 1 : 0;
  }
  view_.pred_off[V] = static_cast<std::uint32_t>(view_.pred_rank.size());
}

// src/graph.cpp:321-337: one column per aligned group, groups are
// contiguous in rank order
std::vector<std::uint32_t> Graph::InitializeMultipleSequenceAlignment(std::uint32_t* row_size) const {
  std::vector<std::uint32_t> column(nodes_.size(), 0);
  std::uint32_t col = 0;
  for (std::size_t r = 0; r < rank_to_node_.size(); ++col) {
    const Node* nd = rank_to_node_[r];
    column[nd->id] = col;
    for (const Node* a : nd->aligned_nodes) column[a->id] = col;
    r += 1 + nd->aligned_nodes.size();
  }
  if (row_size) *row_size = col;
  return column;
}

// src/graph.cpp:339-366
std::vector<std::string> Graph::GenerateMultipleSequenceAlignment(bool include_consensus) {
  std::uint32_t width = 0;
  const auto column = InitializeMultipleSequenceAlignment(&width);
  std::vector<std::string> rows;
  rows.reserve(sequences_.size() + (include_consensus ? 1 : 0));
  for (std::uint32_t s = 0; s < sequences_.size(); ++s) {
    std::string row(width, '-');
    for (const Node* nd = sequences_[s]; nd; nd = nd->Successor(s)) {
      row[column[nd->id]] = static_cast<char>(decoder_[nd->code]);
    }
    rows.push_back(std::move(row));
  }
  if (include_consensus) {
    TraverseHeaviestBundle();
    std::string row(width, '-');
    for (const Node* nd : consensus_) row[column[nd->id]] = static_cast<char>(decoder_[nd->code]);
    rows.push_back(std::move(row));
  }
  return rows;
}

// src/graph.cpp:368-375
std::string Graph::GenerateConsensus() {
  TraverseHeaviestBundle();
  std::string out;
  out.reserve(consensus_.size());
  for (const Node* nd : consensus_) out.push_back(static_cast<char>(decoder_[nd->code]));
  return out;
}

// src/graph.cpp:433-476: heaviest bundle -- each node keeps its heaviest
// in-edge (ties go to the later edge whose tail scores at least as high),
// the best node is extended to a sink by branch completion, then the path
// is read back through the chosen predecessors.
void Graph::TraverseHeaviestBundle() {
  if (rank_to_node_.empty()) return;
  std::vector<Node*> pred(nodes_.size(), nullptr);
  std::vector<std::int64_t> score(nodes_.size(), -1);
  Node* best = nullptr;
  for (Node* nd : rank_to_node_) {
    for (Edge* e : nd->inedges) {
      if (score[nd->id] < e->weight ||
          (score[nd->id] == e->weight && score[pred[nd->id]->id] <= score[e->tail->id])) {
        score[nd->id] = e->weight;
        pred[nd->id] = e->tail;
      }
    }
    if (pred[nd->id]) score[nd->id] += score[pred[nd->id]->id];
    if (!best || score[best->id] < score[nd->id]) best = nd;
  }
  if (!best->outedges.empty()) {
    std::vector<std::uint32_t> rank_of(nodes_.size(), 0);
    for (std::uint32_t r = 0; r < rank_to_node_.size(); ++r) rank_of[rank_to_node_[r]->id] = r;
    while (!best->outedges.empty()) best = BranchCompletion(rank_of[best->id], &score, &pred);
  }
  consensus_.clear();
  for (Node* nd = best; nd; nd = pred[nd->id]) consensus_.push_back(nd);
  std::reverse(consensus_.begin(), consensus_.end());
}

// src/graph.cpp:478-516: rescore everything ranked after `rank`, ignoring
// in-edges from nodes that compete with the start node for its successors.
Graph::Node* Graph::BranchCompletion(std::uint32_t rank, std::vector<std::int64_t>* scores,
                                     std::vector<Node*>* predecessors) {
  auto& score = *scores;
  auto& pred = *predecessors;
  Node* start = rank_to_node_[rank];
  for (Edge* e : start->outedges) {
    for (Edge* f : e->head->inedges) {
      if (f->tail != start) score[f->tail->id] = -1;
    }
  }
  Node* best = nullptr;
  for (std::uint32_t r = rank + 1; r < rank_to_node_.size(); ++r) {
    Node* nd = rank_to_node_[r];
    score[nd->id] = -1;
    pred[nd->id] = nullptr;
    for (Edge* e : nd->inedges) {
      if (score[e->tail->id] == -1) continue;
      if (score[nd->id] < e->weight ||
          (score[nd->id] == e->weight && score[pred[nd->id]->id] <= score[e->tail->id])) {
        score[nd->id] = e->weight;
        pred[nd->id] = e->tail;
      }
    }
    if (pred[nd->id]) score[nd->id] += score[pred[nd->id]->id];
    if (!best || score[best->id] < score[nd->id]) best = nd;
  }
  return best;
}

// src/graph.cpp:377-431: the consensus and, per consensus position, its
// coverage -- the reads through the node and its aligned nodes, or with
// verbose one count per code (and a last row of gaps) of the reads whose
// MSA column meets that position.
std::string Graph::GenerateConsensus(std::vector<std::uint32_t>* summary, bool verbose) {
  if (!summary) {
    throw std::invalid_argument("[spoa::Graph::GenerateConsensus] error: invalid ptr to summary");
  }
  std::string dst = GenerateConsensus();
  summary->clear();
  const std::size_t C = consensus_.size();
  if (!verbose) {
    for (const Node* nd : consensus_) {
      std::uint32_t cov = nd->Coverage();
      for (const Node* a : nd->aligned_nodes) cov += a->Coverage();
      summary->push_back(cov);
    }
    return dst;
  }
  summary->assign((num_codes_ + 1) * C, 0);
  const std::vector<std::uint32_t> col = InitializeMultipleSequenceAlignment();
  for (std::uint32_t s = 0; s < sequences_.size(); ++s) {
    const Node* nd = sequences_[s];
    std::uint32_t c = 0, last = 0, column = col[nd->id];
    bool seen = false;  // an earlier node of this read met the consensus
    while (true) {
      // the first consensus position at or after the read's column
      while (c < C && col[consensus_[c]->id] < column) ++c;
      if (c < C && col[consensus_[c]->id] == column) {
        if (seen) {
          for (std::uint32_t j = last + 1; j < c; ++j) ++(*summary)[num_codes_ * C + j];
        }
        seen = true;
        last = c;
        ++(*summary)[nd->code * C + c];
      }
      if (c == C || !(nd = nd->Successor(s))) break;
      column = col[nd->id];
    }
  }
  return dst;
}

// src/graph.cpp:518-539: the nodes reached backwards from `from` (in-edges
// and aligned nodes) without going below node id `floor`.
std::vector<bool> Graph::ExtractSubgraph(const Node* from, const Node* floor) const {
  std::vector<bool> in(nodes_.size(), false);
  std::vector<const Node*> todo{from};
  while (!todo.empty()) {
    const Node* nd = todo.back();
    todo.pop_back();
    if (in[nd->id] || nd->id < floor->id) continue;
    for (const Edge* e : nd->inedges) todo.push_back(e->tail);
    for (const Node* a : nd->aligned_nodes) todo.push_back(a);
    in[nd->id] = true;
  }
  return in;
}

// src/graph.cpp:541-595: the part of the graph between nodes begin and end
// (everything end reaches backwards, down to id begin) as a graph of its
// own with the same codes; subgraph_to_graph maps its node ids to this
// graph's nodes.  Edges keep their weights (as 32-bit additions, as the
// reference's AddEdge takes them) and carry no sequence labels.
Graph Graph::Subgraph(std::uint32_t begin, std::uint32_t end, std::vector<const Node*>* subgraph_to_graph) const {
  if (!subgraph_to_graph) {
    throw std::invalid_argument("[spoa::Graph::Subgraph] error: invalid ptr to subgraph_to_graph");
  }
  if (begin >= nodes_.size() || end >= nodes_.size()) {
    throw std::invalid_argument("[spoa::Graph::Subgraph] error: node out of range");
  }
  const std::vector<bool> in = ExtractSubgraph(nodes_[end].get(), nodes_[begin].get());
  Graph sub;
  sub.num_codes_ = num_codes_;
  sub.coder_ = coder_;
  sub.decoder_ = decoder_;
  subgraph_to_graph->assign(nodes_.size(), nullptr);
  std::vector<Node*> to_sub(nodes_.size(), nullptr);
  for (const auto& nd : nodes_) {
    if (!in[nd->id]) continue;
    Node* s = sub.AddNode(nd->code);
    to_sub[nd->id] = s;
    (*subgraph_to_graph)[s->id] = nd.get();
  }
  for (const auto& nd : nodes_) {
    if (!in[nd->id]) continue;
    Node* s = to_sub[nd->id];
    for (const Edge* e : nd->inedges) {
      if (to_sub[e->tail->id]) sub.AddEdge(to_sub[e->tail->id], s, static_cast<std::uint32_t>(e->weight));
    }
    for (const Node* a : nd->aligned_nodes) {
      if (to_sub[a->id]) s->aligned_nodes.push_back(to_sub[a->id]);
    }
  }
  sub.TopologicalSort();
  return sub;
}

// src/graph.cpp:597-605: node ids of an alignment against a subgraph back to
// this graph's ids.
void Graph::UpdateAlignment(const std::vector<const Node*>& subgraph_to_graph, Alignment* alignment) const {
  for (auto& pr : *alignment) {
    if (pr.first != -1) pr.first = static_cast<std::int32_t>(subgraph_to_graph[pr.first]->id);
  }
}

// src/graph.cpp:607-647: Graphviz text of the graph -- nodes labelled
// "id - base", out-edges labelled with their weights, aligned pairs as
// dotted undirected lines, the last consensus highlighted.
void Graph::PrintDot(const std::string& path) const {
  if (path.empty()) return;
  std::ofstream os(path);
  std::vector<std::int32_t> on_consensus(nodes_.size(), -1);
  std::int32_t k = 0;
  for (const Node* nd : consensus_) on_consensus[nd->id] = k++;
  os << "digraph " << sequences_.size() << " {" << std::endl << "  graph [rankdir = LR]" << std::endl;
  for (const auto& nd : nodes_) {
    os << "  " << nd->id << "[label = \"" << nd->id << " - " << static_cast<char>(decoder_[nd->code]) << "\"";
    if (on_consensus[nd->id] != -1) os << ", style = filled, fillcolor = goldenrod1";
    os << "]" << std::endl;
    for (const Edge* e : nd->outedges) {
      os << "  " << nd->id << " -> " << e->head->id << " [label = \"" << e->weight << "\"";
      if (on_consensus[nd->id] + 1 == on_consensus[e->head->id]) os << ", color = goldenrod1";
      os << "]" << std::endl;
    }
    for (const Node* a : nd->aligned_nodes) {
      if (a->id > nd->id) os << "  " << nd->id << " -> " << a->id << " [style = dotted, arrowhead = none]" << std::endl;
    }
  }
  os << "}" << std::endl;
}

// src/graph.cpp:649-658
void Graph::Clear() {
  num_codes_ = 0;
  std::fill(coder_.begin(), coder_.end(), -1);
  std::fill(decoder_.begin(), decoder_.end(), -1);
  sequences_.clear();
  nodes_.clear();
  edges_.clear();
  rank_to_node_.clear();
  consensus_.clear();
  view_ = RankView();
  ++version_;
}

}  // namespace spoa
