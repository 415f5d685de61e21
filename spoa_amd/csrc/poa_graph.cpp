// Flat partial-order graph (see poa_graph.hpp).

#include "poa_graph.hpp"

#include <algorithm>
#include <stdexcept>

namespace spoa_amd {

PoaGraph::PoaGraph() { Clear(); }

// src/graph.cpp:649-658
void PoaGraph::Clear() {
  for (int i = 0; i < 256; ++i) coder_[i] = decoder_[i] = -1;
  num_codes_ = 0;
  code_.clear();
  first_in_.clear();
  last_in_.clear();
  first_out_.clear();
  last_out_.clear();
  aligned_.clear();
  tail_.clear();
  head_.clear();
  next_in_.clear();
  next_out_.clear();
  weight_.clear();
  paths_.clear();
  rank_to_node_.clear();
  consensus_.clear();
  RebuildView();
}

// src/graph.cpp:76-79
std::uint32_t PoaGraph::AddNode(std::uint32_t code) {
  code_.push_back(static_cast<std::uint8_t>(code));
  first_in_.push_back(kNil);
  last_in_.push_back(kNil);
  first_out_.push_back(kNil);
  last_out_.push_back(kNil);
  aligned_.emplace_back();
  return num_nodes() - 1;
}

// src/graph.cpp:81-91: a tail->head edge that exists gains the weight
void PoaGraph::AddEdge(std::uint32_t t, std::uint32_t h, std::uint32_t w) {
  for (std::uint32_t e = first_out_[t]; e != kNil; e = next_out_[e]) {
    if (head_[e] == h) {
      weight_[e] += w;
      return;
    }
  }
  const std::uint32_t e = static_cast<std::uint32_t>(tail_.size());
  tail_.push_back(t);
  head_.push_back(h);
  weight_.push_back(w);
  next_in_.push_back(kNil);
  next_out_.push_back(kNil);
  if (last_out_[t] == kNil) {
    first_out_[t] = e;
  } else {
    next_out_[last_out_[t]] = e;
  }
  last_out_[t] = e;
  if (last_in_[h] == kNil) {
    first_in_[h] = e;
  } else {
    next_in_[last_in_[h]] = e;
  }
  last_in_[h] = e;
}

// src/graph.cpp:93-110: a fresh chain for seq[begin, end); returns its
// first node or -1
std::int64_t PoaGraph::AddSequence(const char* seq, const std::uint32_t* w, std::uint32_t begin, std::uint32_t end,
                                   std::vector<std::uint32_t>* path) {
  if (begin == end) return -1;
  std::int64_t first = -1, prev = -1;
  for (std::uint32_t i = begin; i < end; ++i) {
    const std::uint32_t curr = AddNode(static_cast<std::uint32_t>(coder_[static_cast<std::uint8_t>(seq[i])]));
    if (prev >= 0) AddEdge(static_cast<std::uint32_t>(prev), curr, w[i - 1] + w[i]);
    if (first < 0) first = curr;
    prev = curr;
    path->push_back(curr);
  }
  return first;
}

// src/graph.cpp:155-247
void PoaGraph::AddAlignment(const std::int32_t* pairs, std::uint32_t n_pairs, const char* seq, std::uint32_t len,
                            const std::uint32_t* w) {
  if (len == 0) return;
  for (std::uint32_t i = 0; i < len; ++i) {
    const std::uint8_t ch = static_cast<std::uint8_t>(seq[i]);
    if (coder_[ch] == -1) {
      coder_[ch] = static_cast<std::int32_t>(num_codes_);
      decoder_[num_codes_++] = static_cast<std::uint8_t>(seq[i]);
    }
  }
  std::vector<std::uint32_t> path;
  if (n_pairs == 0) {
    AddSequence(seq, w, 0, len, &path);
    paths_.push_back(std::move(path));
    TopologicalSort();
    return;
  }
  std::int32_t first_pos = -1, last_pos = -1;
  for (std::uint32_t k = 0; k < n_pairs; ++k) {
    const std::int32_t pos = pairs[2 * k + 1];
    if (pos == -1) continue;
    if (pos < 0 || pos >= static_cast<std::int32_t>(len)) {
      throw std::invalid_argument("[spoa::Graph::AddAlignment] error: invalid alignment");
    }
    if (first_pos == -1) first_pos = pos;
    last_pos = pos;
  }
  if (first_pos == -1) {
    throw std::invalid_argument("[spoa::Graph::AddAlignment] error: missing sequence in alignment");
  }
  // unaligned prefix and suffix become fresh chains (the suffix is created
  // before the aligned nodes but visited after them)
  const std::int64_t begin0 = AddSequence(seq, w, 0, static_cast<std::uint32_t>(first_pos), &path);
  std::int64_t prev = begin0 >= 0 ? static_cast<std::int64_t>(num_nodes()) - 1 : -1;
  std::vector<std::uint32_t> suffix;
  const std::int64_t last = AddSequence(seq, w, static_cast<std::uint32_t>(last_pos) + 1, len, &suffix);
  for (std::uint32_t k = 0; k < n_pairs; ++k) {
    const std::int32_t id = pairs[2 * k], pos = pairs[2 * k + 1];
    if (pos == -1) continue;
    const std::uint32_t c = static_cast<std::uint32_t>(coder_[static_cast<std::uint8_t>(seq[pos])]);
    std::uint32_t curr;
    if (id == -1) {
      curr = AddNode(c);
    } else {
      const std::uint32_t jt = static_cast<std::uint32_t>(id);
      if (code_[jt] == c) {
        curr = jt;
      } else {
        curr = kNil;
        for (std::uint32_t kt : aligned_[jt]) {
          if (code_[kt] == c) {
            curr = kt;
            break;
          }
        }
        if (curr == kNil) {
          curr = AddNode(c);
          const std::vector<std::uint32_t> group = aligned_[jt];
          for (std::uint32_t kt : group) {
            aligned_[kt].push_back(curr);
            aligned_[curr].push_back(kt);
          }
          aligned_[jt].push_back(curr);
          aligned_[curr].push_back(jt);
        }
      }
    }
    if (prev >= 0) AddEdge(static_cast<std::uint32_t>(prev), curr, w[pos - 1] + w[pos]);
    prev = curr;
    path.push_back(curr);
  }
  if (last >= 0) {
    AddEdge(static_cast<std::uint32_t>(prev), static_cast<std::uint32_t>(last), w[last_pos] + w[last_pos + 1]);
  }
  path.insert(path.end(), suffix.begin(), suffix.end());
  paths_.push_back(std::move(path));
  TopologicalSort();
}

// src/graph.cpp:249-303: depth-first over node ids; a node is finished once
// its in-edge tails (and, for a group's first member, its aligned nodes)
// are; the first member is ranked with its whole group.
void PoaGraph::TopologicalSort() {
  const std::uint32_t V = num_nodes();
  rank_to_node_.clear();
  marks_.assign(V, 0);
  ignored_.assign(V, 0);
  stack_.clear();
  // the kernel view is built as nodes are ranked: every predecessor of a
  // node is ranked before it (the order is topological)
  rank_of_.resize(V);
  v_char_.clear();
  v_sink_.clear();
  v_id_.clear();
  v_off_.clear();
  v_pred_.clear();
  auto emit = [&](std::uint32_t id) {
    rank_of_[id] = static_cast<std::uint32_t>(rank_to_node_.size());
    rank_to_node_.push_back(id);
    v_off_.push_back(static_cast<std::uint32_t>(v_pred_.size()));
    for (std::uint32_t e = first_in_[id]; e != kNil; e = next_in_[e]) v_pred_.push_back(rank_of_[tail_[e]]);
    v_char_.push_back(static_cast<std::uint8_t>(decoder_[code_[id]]));
    v_id_.push_back(static_cast<std::int32_t>(id));
    v_sink_.push_back(first_out_[id] == kNil ? 1 : 0);
  };
  for (std::uint32_t root = 0; root < V; ++root) {
    if (marks_[root] != 0) continue;
    stack_.push_back(root);
    while (!stack_.empty()) {
      const std::uint32_t curr = stack_.back();
      bool valid = true;
      if (marks_[curr] != 2) {
        for (std::uint32_t e = first_in_[curr]; e != kNil; e = next_in_[e]) {
          const std::uint32_t t = tail_[e];
          if (marks_[t] != 2) {
            stack_.push_back(t);
            valid = false;
          }
        }
        if (!ignored_[curr]) {
          for (std::uint32_t a : aligned_[curr]) {
            if (marks_[a] != 2) {
              stack_.push_back(a);
              ignored_[a] = 1;
              valid = false;
            }
          }
        }
        if (valid) {
          marks_[curr] = 2;
          if (!ignored_[curr]) {
            emit(curr);
            for (std::uint32_t a : aligned_[curr]) emit(a);
          }
        } else {
          marks_[curr] = 1;
        }
      }
      if (valid) stack_.pop_back();
    }
  }
  v_off_.push_back(static_cast<std::uint32_t>(v_pred_.size()));
}

void PoaGraph::RebuildView() {
  const std::uint32_t V = static_cast<std::uint32_t>(rank_to_node_.size());
  rank_of_.resize(num_nodes());
  for (std::uint32_t r = 0; r < V; ++r) rank_of_[rank_to_node_[r]] = r;
  v_char_.resize(V);
  v_sink_.resize(V);
  v_id_.resize(V);
  v_off_.resize(V + 1);
  v_pred_.clear();
  v_pred_.reserve(tail_.size());
  for (std::uint32_t r = 0; r < V; ++r) {
    const std::uint32_t id = rank_to_node_[r];
    v_off_[r] = static_cast<std::uint32_t>(v_pred_.size());
    for (std::uint32_t e = first_in_[id]; e != kNil; e = next_in_[e]) v_pred_.push_back(rank_of_[tail_[e]]);
    v_char_[r] = static_cast<std::uint8_t>(decoder_[code_[id]]);
    v_id_[r] = static_cast<std::int32_t>(id);
    v_sink_[r] = first_out_[id] == kNil ? 1 : 0;
  }
  v_off_[V] = static_cast<std::uint32_t>(v_pred_.size());
}

spoa_hip_graph PoaGraph::View() const {
  spoa_hip_graph g{};
  g.num_nodes = static_cast<std::uint32_t>(rank_to_node_.size());
  g.row_char = v_char_.data();
  g.node_id = v_id_.data();
  g.pred_off = v_off_.data();
  g.pred_rank = v_pred_.data();
  g.is_sink = v_sink_.data();
  return g;
}

// src/graph.cpp:478-516
std::uint32_t PoaGraph::BranchCompletion(std::uint32_t rank, std::vector<std::int64_t>& score,
                                         std::vector<std::uint32_t>& pred) {
  const std::uint32_t start = rank_to_node_[rank];
  for (std::uint32_t e = first_out_[start]; e != kNil; e = next_out_[e]) {
    for (std::uint32_t f = first_in_[head_[e]]; f != kNil; f = next_in_[f]) {
      if (tail_[f] != start) score[tail_[f]] = -1;
    }
  }
  std::uint32_t best = kNil;
  for (std::uint32_t i = rank + 1; i < rank_to_node_.size(); ++i) {
    const std::uint32_t it = rank_to_node_[i];
    score[it] = -1;
    pred[it] = kNil;
    for (std::uint32_t e = first_in_[it]; e != kNil; e = next_in_[e]) {
      const std::uint32_t t = tail_[e];
      if (score[t] == -1) continue;
      if (score[it] < weight_[e] || (score[it] == weight_[e] && score[pred[it]] <= score[t])) {
        score[it] = weight_[e];
        pred[it] = t;
      }
    }
    if (pred[it] != kNil) score[it] += score[pred[it]];
    if (best == kNil || score[best] < score[it]) best = it;
  }
  return best;
}

// src/graph.cpp:433-476
void PoaGraph::TraverseHeaviestBundle() {
  if (rank_to_node_.empty()) return;
  std::vector<std::uint32_t> pred(num_nodes(), kNil);
  std::vector<std::int64_t> score(num_nodes(), -1);
  std::uint32_t best = kNil;
  for (std::uint32_t it : rank_to_node_) {
    for (std::uint32_t e = first_in_[it]; e != kNil; e = next_in_[e]) {
      const std::uint32_t t = tail_[e];
      if (score[it] < weight_[e] || (score[it] == weight_[e] && score[pred[it]] <= score[t])) {
        score[it] = weight_[e];
        pred[it] = t;
      }
    }
    if (pred[it] != kNil) score[it] += score[pred[it]];
    if (best == kNil || score[best] < score[it]) best = it;
  }
  while (first_out_[best] != kNil) best = BranchCompletion(rank_of_[best], score, pred);
  consensus_.clear();
  for (std::uint32_t v = best; v != kNil; v = pred[v]) consensus_.push_back(v);
  std::reverse(consensus_.begin(), consensus_.end());
}

// src/graph.cpp:368-375
std::string PoaGraph::GenerateConsensus() {
  TraverseHeaviestBundle();
  std::string s;
  s.reserve(consensus_.size());
  for (std::uint32_t v : consensus_) s.push_back(static_cast<char>(decoder_[code_[v]]));
  return s;
}

// src/graph.cpp:321-337: one column per aligned group (groups are
// contiguous in rank order)
std::vector<std::uint32_t> PoaGraph::MsaColumns(std::uint32_t* width) const {
  std::vector<std::uint32_t> col(num_nodes(), 0);
  std::uint32_t j = 0;
  for (std::uint32_t i = 0; i < rank_to_node_.size(); ++i, ++j) {
    const std::uint32_t it = rank_to_node_[i];
    col[it] = j;
    for (std::uint32_t a : aligned_[it]) {
      col[a] = j;
      ++i;
    }
  }
  *width = j;
  return col;
}

// src/graph.cpp:339-366
std::vector<std::string> PoaGraph::GenerateMultipleSequenceAlignment(bool include_consensus) {
  std::uint32_t width = 0;
  const std::vector<std::uint32_t> col = MsaColumns(&width);
  std::vector<std::string> rows;
  rows.reserve(paths_.size() + (include_consensus ? 1 : 0));
  for (const auto& path : paths_) {
    std::string row(width, '-');
    for (std::uint32_t v : path) row[col[v]] = static_cast<char>(decoder_[code_[v]]);
    rows.push_back(std::move(row));
  }
  if (include_consensus) {
    TraverseHeaviestBundle();
    std::string row(width, '-');
    for (std::uint32_t v : consensus_) row[col[v]] = static_cast<char>(decoder_[code_[v]]);
    rows.push_back(std::move(row));
  }
  return rows;
}

}  // namespace spoa_amd
