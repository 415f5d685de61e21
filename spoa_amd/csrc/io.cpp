// Sequence files and GFA output (see io.hpp).

#include "io.hpp"

#include <zlib.h>

#include <algorithm>
#include <stdexcept>

namespace spoa_amd {

namespace {

bool EndsWith(const std::string& s, const std::string& suffix) {
  return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

std::string RightStrip(std::string s) {
  while (!s.empty() && (s.back() == ' ' || s.back() == '\t' || s.back() == '\r' || s.back() == '\n')) s.pop_back();
  return s;
}

// All lines of a gzip or plain text file.
std::vector<std::string> ReadLines(const std::string& path) {
  gzFile f = gzopen(path.c_str(), "rb");
  if (!f) throw std::invalid_argument("[spoa::CreateParser] error: unable to open file " + path);
  std::vector<std::string> lines;
  std::string cur;
  char buf[1 << 16];
  int n;
  while ((n = gzread(f, buf, sizeof(buf))) > 0) {
    for (int i = 0; i < n; ++i) {
      if (buf[i] == '\n') {
        lines.push_back(RightStrip(cur));
        cur.clear();
      } else {
        cur.push_back(buf[i]);
      }
    }
  }
  const bool failed = n < 0;
  gzclose(f);
  if (failed) throw std::invalid_argument("[spoa::CreateParser] error: unable to read file " + path);
  if (!cur.empty()) lines.push_back(RightStrip(cur));
  return lines;
}

}  // namespace

bool SequenceFormat(const std::string& path, bool* fastq) {
  for (const char* ext : {".fasta", ".fasta.gz", ".fna", ".fna.gz", ".faa", ".faa.gz", ".fa", ".fa.gz"}) {
    if (EndsWith(path, ext)) {
      *fastq = false;
      return true;
    }
  }
  for (const char* ext : {".fastq", ".fastq.gz", ".fq", ".fq.gz"}) {
    if (EndsWith(path, ext)) {
      *fastq = true;
      return true;
    }
  }
  return false;
}

std::vector<SeqRecord> ReadSequences(const std::string& path) {
  bool fastq = false;
  if (!SequenceFormat(path, &fastq)) {
    throw std::invalid_argument(
        "[spoa::CreateParser] error: file " + path +
        " has unsupported format extension (valid extensions: .fasta, .fasta.gz, .fna, .fna.gz, .faa, .faa.gz, "
        ".fa, .fa.gz, .fastq, .fastq.gz, .fq, .fq.gz)");
  }
  const std::vector<std::string> lines = ReadLines(path);
  std::vector<SeqRecord> out;
  if (!fastq) {
    for (const std::string& l : lines) {
      if (l.empty()) continue;
      if (l[0] == '>') {
        out.push_back(SeqRecord{l.substr(1), "", ""});
      } else {
        if (out.empty()) throw std::invalid_argument("[spoa::FastaParser] error: invalid file format");
        out.back().data += l;
      }
    }
    return out;
  }
  std::size_t i = 0;
  while (i < lines.size()) {
    if (lines[i].empty()) {
      ++i;
      continue;
    }
    if (lines[i][0] != '@' || i + 3 >= lines.size() || lines[i + 2].empty() || lines[i + 2][0] != '+') {
      throw std::invalid_argument("[spoa::FastqParser] error: invalid file format");
    }
    SeqRecord r{lines[i].substr(1), lines[i + 1], lines[i + 3]};
    if (r.quality.size() != r.data.size()) {
      throw std::invalid_argument("[spoa::FastqParser] error: sequence and quality lengths differ");
    }
    out.push_back(std::move(r));
    i += 4;
  }
  return out;
}

bool WriteGfa(std::ostream& os, const spoa::Graph& graph, const std::vector<std::string>& headers,
              const std::vector<bool>& is_reversed, bool include_consensus) {
  const auto& seqs = graph.sequences();
  if (headers.size() < seqs.size()) return false;
  if (!is_reversed.empty() && is_reversed.size() < seqs.size()) return false;
  std::vector<bool> on_consensus(graph.nodes().size(), false);
  for (const auto* nd : graph.consensus()) on_consensus[nd->id] = true;
  os << "H\tVN:Z:1.0" << "\n";
  for (const auto& nd : graph.nodes()) {
    os << "S\t" << nd->id + 1 << "\t" << static_cast<char>(graph.decoder(static_cast<std::uint8_t>(nd->code)));
    if (on_consensus[nd->id]) os << "\tic:Z:true";
    os << "\n";
    for (const auto* e : nd->outedges) {
      os << "L\t" << nd->id + 1 << "\t+\t" << e->head->id + 1 << "\t+\tOM\tew:f:" << e->weight;
      if (on_consensus[nd->id] && on_consensus[e->head->id]) os << "\tic:Z:true";
      os << "\n";
    }
  }
  for (std::uint32_t s = 0; s < seqs.size(); ++s) {
    os << "P\t" << headers[s] << "\t";
    std::vector<std::uint32_t> path;
    for (const spoa::Graph::Node* nd = seqs[s]; nd; nd = nd->Successor(s)) path.push_back(nd->id + 1);
    const bool rev = !is_reversed.empty() && is_reversed[s];
    if (rev) std::reverse(path.begin(), path.end());
    for (std::size_t j = 0; j < path.size(); ++j) os << (j ? "," : "") << path[j] << (rev ? "-" : "+");
    os << "\t*" << "\n";
  }
  if (include_consensus) {
    os << "P\tConsensus\t";
    const auto& cons = graph.consensus();
    for (std::size_t j = 0; j < cons.size(); ++j) os << (j ? "," : "") << cons[j]->id + 1 << "+";
    os << "\t*" << "\n";
  }
  return true;
}

}  // namespace spoa_amd
