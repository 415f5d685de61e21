// spoa command-line tool on the MI355X engine: the reference tool's options
// and outputs (src/main.cpp:15-23, 61-117, 203-352) -- FASTA / FASTQ input
// (optionally gzip-compressed), reads aligned in file order (with -s both
// strands, the better kept), consensus / MSA (FASTA) / GFA on stdout, the
// graph in DOT with -d.  Alignments run on the GPU through
// spoa::AlignmentEngine; without a usable device the tool fails.

#include <getopt.h>

#include <algorithm>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "batch.hpp"
#include "io.hpp"
#include "spoa/spoa.hpp"

namespace {

const char kVersion[] = "spoa_amd 4.0.8";

void Help() {
  std::cout << "usage: spoa [options ...] <sequences>\n"
               "\n"
               "  # default output is stdout\n"
               "  <sequences>\n"
               "    input file in FASTA/FASTQ format (can be compressed with gzip)\n"
               "\n"
               "  options:\n"
               "    -m <int>\n      default: 5\n      score for matching bases\n"
               "    -n <int>\n      default: -4\n      score for mismatching bases\n"
               "    -g <int>\n      default: -8\n      gap opening penalty (must be non-positive)\n"
               "    -e <int>\n      default: -6\n      gap extension penalty (must be non-positive)\n"
               "    -q <int>\n      default: -10\n      gap opening penalty of the second affine function\n"
               "      (must be non-positive)\n"
               "    -c <int>\n      default: -4\n      gap extension penalty of the second affine function\n"
               "      (must be non-positive)\n"
               "    -l, --algorithm <int>\n      default: 0\n      alignment mode:\n"
               "        0 - local (Smith-Waterman)\n        1 - global (Needleman-Wunsch)\n"
               "        2 - semi-global\n"
               "    -r, --result <int> (option can be used multiple times)\n      default: 0\n"
               "      result mode:\n        0 - consensus (FASTA)\n"
               "        1 - multiple sequence alignment (FASTA)\n        2 - 0 & 1 (FASTA)\n"
               "        3 - partial order graph (GFA)\n        4 - 0 & 3 (GFA)\n"
               "    -d, --dot <file>\n      output file for the partial order graph in DOT format\n"
               "    -s, --strand-ambiguous\n      for each sequence pick the strand with the better alignment\n"
               "    --version\n      prints the version number\n"
               "    -h, --help\n      prints the usage\n"
               "\n"
               "  gap mode:\n    linear if g >= e\n    affine if g <= q or e >= c\n"
               "    convex otherwise (default)\n";
}

}  // namespace

int main(int argc, char** argv) {
  static struct option options[] = {{"algorithm", required_argument, nullptr, 'l'},
                                    {"result", required_argument, nullptr, 'r'},
                                    {"dot", required_argument, nullptr, 'd'},
                                    {"strand-ambiguous", no_argument, nullptr, 's'},
                                    {"version", no_argument, nullptr, 'v'},
                                    {"help", no_argument, nullptr, 'h'},
                                    {nullptr, 0, nullptr, 0}};
  std::int8_t m = 5, n = -4, g = -8, e = -6, q = -10, c = -4;
  std::uint8_t algorithm = 0;
  std::vector<std::uint8_t> results = {0};
  std::string dot_path;
  bool strand = false;
  int opt;
  while ((opt = getopt_long(argc, argv, "m:n:g:e:q:c:l:r:d:sh", options, nullptr)) != -1) {
    switch (opt) {
      case 'm': m = static_cast<std::int8_t>(std::atoi(optarg)); break;
      case 'n': n = static_cast<std::int8_t>(std::atoi(optarg)); break;
      case 'g': g = static_cast<std::int8_t>(std::atoi(optarg)); break;
      case 'e': e = static_cast<std::int8_t>(std::atoi(optarg)); break;
      case 'q': q = static_cast<std::int8_t>(std::atoi(optarg)); break;
      case 'c': c = static_cast<std::int8_t>(std::atoi(optarg)); break;
      case 'l': algorithm = static_cast<std::uint8_t>(std::atoi(optarg)); break;
      case 'r': results.push_back(static_cast<std::uint8_t>(std::atoi(optarg))); break;
      case 'd': dot_path = optarg; break;
      case 's': strand = true; break;
      case 'v': std::cout << kVersion << std::endl; return 0;
      case 'h': Help(); return 0;
      default: return 1;
    }
  }
  if (results.size() > 1) results.erase(results.begin());
  if (optind >= argc) {
    std::cerr << "[spoa::] error: missing input file!" << std::endl;
    Help();
    return 1;
  }
  std::vector<spoa_amd::SeqRecord> seqs;
  std::unique_ptr<spoa::AlignmentEngine> engine;
  try {
    seqs = spoa_amd::ReadSequences(argv[optind]);
    engine = spoa::AlignmentEngine::Create(static_cast<spoa::AlignmentType>(algorithm), m, n, g, e, q, c);
    std::size_t max_len = 0;
    for (const auto& s : seqs) max_len = std::max(max_len, s.data.size());
    engine->Prealloc(static_cast<std::uint32_t>(max_len), 4);
  } catch (const std::exception& ex) {
    std::cerr << ex.what() << std::endl;
    return 1;
  }
  spoa::Graph graph;
  std::vector<bool> is_reversed;
  for (auto& s : seqs) {
    try {
      std::int32_t score = 0;
      spoa::Alignment alignment = engine->Align(s.data, graph, &score);
      if (strand) {
        const std::string rc = spoa_amd::ReverseComplement(s.data.data(), s.data.size());
        std::int32_t score_rev = 0;
        spoa::Alignment alignment_rev = engine->Align(rc, graph, &score_rev);
        if (score >= score_rev) {
          is_reversed.push_back(false);
        } else {
          alignment = std::move(alignment_rev);
          s.data = rc;
          std::reverse(s.quality.begin(), s.quality.end());
          is_reversed.push_back(true);
        }
      }
      if (s.quality.empty()) {
        graph.AddAlignment(alignment, s.data);
      } else {
        graph.AddAlignment(alignment, s.data, s.quality);
      }
    } catch (const std::exception& ex) {
      std::cerr << ex.what() << std::endl;
      return 1;
    }
  }
  for (std::uint8_t r : results) {
    switch (r) {
      case 0: {
        const std::string cons = graph.GenerateConsensus();
        std::cout << ">Consensus LN:i:" << cons.size() << "\n" << cons << "\n";
        break;
      }
      case 1:
      case 2: {
        const auto msa = graph.GenerateMultipleSequenceAlignment(r == 2);
        for (std::size_t i = 0; i < msa.size(); ++i) {
          std::cout << ">" << (i < seqs.size() ? seqs[i].name : std::string("Consensus")) << "\n" << msa[i] << "\n";
        }
        break;
      }
      case 3:
      case 4: {
        std::vector<std::string> headers;
        for (const auto& s : seqs) headers.push_back(s.name);
        graph.GenerateConsensus();
        if (!spoa_amd::WriteGfa(std::cout, graph, headers, is_reversed, r == 4)) {
          std::cerr << "[spoa::PrintGfa] error: missing header(s)" << std::endl;
        }
        break;
      }
      default:
        break;
    }
  }
  std::cout.flush();
  graph.PrintDot(dot_path);
  return 0;
}
