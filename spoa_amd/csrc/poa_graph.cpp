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
  in_.clear();
  out_.clear();
  aligned_.clear();
  weight_.clear();
  path_.clear();
  path_off_.assign(1, 0);
  rank_to_node_.clear();
  consensus_.clear();
  rank_of_.clear();
  v_char_.clear();
  v_sink_.clear();
  v_id_.clear();
  v_pred_.clear();
  v_off_.assign(1, 0);
}

// src/graph.cpp:76-79
std::uint32_t PoaGraph::AddNode(std::uint32_t code) {
  code_.push_back(static_cast<std::uint8_t>(code));
  in_.add_owner();
  out_.add_owner();
  aligned_.add_owner();
  return num_nodes() - 1;
}

// src/graph.cpp:81-91: a tail->head edge that exists gains the weight
void PoaGraph::AddEdge(std::uint32_t t, std::uint32_t h, std::uint32_t w) {
  for (const Adj* e = out_.begin(t); e != out_.end(t); ++e) {
    if (e->node == h) {
      weight_[e->edge] += w;
      return;
    }
  }
  const std::uint32_t e = static_cast<std::uint32_t>(weight_.size());
  weight_.push_back(w);
  out_.push(t, Adj{h, e});
  in_.push(h, Adj{t, e});
}

// src/graph.cpp:93-110: a fresh chain for seq[begin, end) (consecutive node
// ids); returns its first node or -1
std::int64_t PoaGraph::AddSequence(const char* seq, const std::uint32_t* w, std::uint32_t begin, std::uint32_t end,
                                   bool on_path) {
  if (begin == end) return -1;
  const std::int64_t first = num_nodes();
  for (std::uint32_t i = begin; i < end; ++i) {
    const std::uint32_t curr = AddNode(static_cast<std::uint32_t>(coder_[static_cast<std::uint8_t>(seq[i])]));
    if (i > begin) AddEdge(curr - 1, curr, w[i - 1] + w[i]);
    if (on_path) path_.push_back(curr);
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
  if (n_pairs == 0) {
    AddSequence(seq, w, 0, len, true);
    path_off_.push_back(static_cast<std::uint32_t>(path_.size()));
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
  const std::int64_t begin0 = AddSequence(seq, w, 0, static_cast<std::uint32_t>(first_pos), true);
  std::int64_t prev = begin0 >= 0 ? static_cast<std::int64_t>(num_nodes()) - 1 : -1;
  const std::uint32_t suffix_len = len - static_cast<std::uint32_t>(last_pos) - 1;
  const std::int64_t last = AddSequence(seq, w, static_cast<std::uint32_t>(last_pos) + 1, len, false);
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
        for (const std::uint32_t* a = aligned_.begin(jt); a != aligned_.end(jt); ++a) {
          if (code_[*a] == c) {
            curr = *a;
            break;
          }
        }
        if (curr == kNil) {
          curr = AddNode(c);
          // jt's list is read by index: pushes to other lists may move the pool
          const std::uint32_t group = aligned_.size(jt);
          for (std::uint32_t i = 0; i < group; ++i) {
            const std::uint32_t kt = aligned_.at(jt, i);
            aligned_.push(kt, curr);
            aligned_.push(curr, kt);
          }
          aligned_.push(jt, curr);
          aligned_.push(curr, jt);
        }
      }
    }
    if (prev >= 0) AddEdge(static_cast<std::uint32_t>(prev), curr, w[pos - 1] + w[pos]);
    prev = curr;
    path_.push_back(curr);
  }
  if (last >= 0) {
    AddEdge(static_cast<std::uint32_t>(prev), static_cast<std::uint32_t>(last), w[last_pos] + w[last_pos + 1]);
    for (std::uint32_t i = 0; i < suffix_len; ++i) path_.push_back(static_cast<std::uint32_t>(last) + i);
  }
  path_off_.push_back(static_cast<std::uint32_t>(path_.size()));
  TopologicalSort();
}

// src/graph.cpp:249-303: depth-first over node ids; a node is finished once
// its in-edge tails (and, for a group's first member, its aligned nodes)
// are; the first member is ranked with its whole group.  The kernel view
// is written as nodes are ranked (every predecessor of a node is ranked
// before it).
void PoaGraph::TopologicalSort() {
  const std::uint32_t V = num_nodes();
  // per node: 1 = visited, 2 = finished, 4 = ranked with its group's first member
  marks_.assign(V, 0);
  rank_to_node_.resize(V);
  rank_of_.resize(V);
  v_char_.resize(V);
  v_sink_.resize(V);
  v_id_.resize(V);
  v_off_.resize(V + 1);
  v_pred_.resize(weight_.size());
  std::uint32_t* const r2n = rank_to_node_.data();
  std::uint32_t* const rof = rank_of_.data();
  std::uint32_t* const off = v_off_.data();
  std::uint32_t* const pred = v_pred_.data();
  std::uint8_t* const vch = v_char_.data();
  std::uint8_t* const vsk = v_sink_.data();
  std::int32_t* const vid = v_id_.data();
  std::uint8_t* const mk = marks_.data();
  const std::uint8_t* const cd = code_.data();
  std::uint32_t r = 0, p = 0;
  auto emit = [&](std::uint32_t id) {
    rof[id] = r;
    r2n[r] = id;
    off[r] = p;
    for (const Adj* e = in_.begin(id); e != in_.end(id); ++e) pred[p++] = rof[e->node];
    vch[r] = static_cast<std::uint8_t>(decoder_[cd[id]]);
    vid[r] = static_cast<std::int32_t>(id);
    vsk[r] = out_.size(id) == 0 ? 1 : 0;
    ++r;
  };
  stack_.clear();
  for (std::uint32_t root = 0; root < V; ++root) {
    if (mk[root] != 0) continue;
    stack_.push_back(root);
    while (!stack_.empty()) {
      const std::uint32_t curr = stack_.back();
      if ((mk[curr] & 3) == 1) {
        // Back on top after its first visit: every node that visit pushed
        // (its unfinished in-edge tails and aligned nodes) has been finished
        // and popped above it, so the reference's second scan of its lists
        // would find them all finished.
        mk[curr] |= 2;
        if ((mk[curr] & 4) == 0) {
          emit(curr);
          for (const std::uint32_t* a = aligned_.begin(curr); a != aligned_.end(curr); ++a) emit(*a);
        }
      } else if ((mk[curr] & 2) == 0) {
        bool valid = true;
        for (const Adj* e = in_.begin(curr); e != in_.end(curr); ++e) {
          const std::uint32_t t = e->node;
          if ((mk[t] & 2) == 0) {
            stack_.push_back(t);
            valid = false;
          }
        }
        const bool ignored = (mk[curr] & 4) != 0;
        if (!ignored) {
          for (const std::uint32_t* a = aligned_.begin(curr); a != aligned_.end(curr); ++a) {
            if ((mk[*a] & 2) == 0) {
              stack_.push_back(*a);
              mk[*a] |= 4;
              valid = false;
            }
          }
        }
        if (!valid) {
          mk[curr] |= 1;
          continue;
        }
        mk[curr] |= 2;
        if (!ignored) {
          emit(curr);
          for (const std::uint32_t* a = aligned_.begin(curr); a != aligned_.end(curr); ++a) emit(*a);
        }
      }
      stack_.pop_back();
    }
  }
  off[V] = p;
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
  for (const Adj* e = out_.begin(start); e != out_.end(start); ++e) {
    for (const Adj* f = in_.begin(e->node); f != in_.end(e->node); ++f) {
      if (f->node != start) score[f->node] = -1;
    }
  }
  std::uint32_t best = kNil;
  for (std::uint32_t i = rank + 1; i < rank_to_node_.size(); ++i) {
    const std::uint32_t it = rank_to_node_[i];
    score[it] = -1;
    pred[it] = kNil;
    for (const Adj* ep = in_.begin(it); ep != in_.end(it); ++ep) {
      const std::uint32_t e = ep->edge, t = ep->node;
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
    for (const Adj* ep = in_.begin(it); ep != in_.end(it); ++ep) {
      const std::uint32_t e = ep->edge, t = ep->node;
      if (score[it] < weight_[e] || (score[it] == weight_[e] && score[pred[it]] <= score[t])) {
        score[it] = weight_[e];
        pred[it] = t;
      }
    }
    if (pred[it] != kNil) score[it] += score[pred[it]];
    if (best == kNil || score[best] < score[it]) best = it;
  }
  while (out_.size(best) != 0) best = BranchCompletion(rank_of_[best], score, pred);
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
    for (const std::uint32_t* a = aligned_.begin(it); a != aligned_.end(it); ++a) {
      col[*a] = j;
      ++i;
    }
  }
  *width = j;
  return col;
}

// src/graph.cpp:339-366
void PoaGraph::MultipleSequenceAlignmentBlock(bool include_consensus, std::string* block, std::uint32_t* n_rows,
                                              std::uint32_t* width) {
  std::uint32_t w = 0;
  const std::vector<std::uint32_t> col = MsaColumns(&w);
  const std::uint32_t n = num_sequences() + (include_consensus ? 1 : 0);
  block->assign(static_cast<std::size_t>(n) * w, '-');
  char* const out = block->empty() ? nullptr : &(*block)[0];
  for (std::uint32_t s = 0; s < num_sequences(); ++s) {
    char* const row = out + static_cast<std::size_t>(s) * w;
    for (std::uint32_t k = path_off_[s]; k < path_off_[s + 1]; ++k) {
      const std::uint32_t v = path_[k];
      row[col[v]] = static_cast<char>(decoder_[code_[v]]);
    }
  }
  if (include_consensus) {
    TraverseHeaviestBundle();
    char* const row = out + static_cast<std::size_t>(num_sequences()) * w;
    for (std::uint32_t v : consensus_) row[col[v]] = static_cast<char>(decoder_[code_[v]]);
  }
  *n_rows = n;
  *width = w;
}

std::vector<std::string> PoaGraph::GenerateMultipleSequenceAlignment(bool include_consensus) {
  std::string block;
  std::uint32_t n = 0, w = 0;
  MultipleSequenceAlignmentBlock(include_consensus, &block, &n, &w);
  std::vector<std::string> rows;
  rows.reserve(n);
  for (std::uint32_t i = 0; i < n; ++i) rows.emplace_back(block, static_cast<std::size_t>(i) * w, w);
  return rows;
}

}  // namespace spoa_amd
