// Sequence files and graph output of the command-line tool (the reference's
// src/main.cpp): FASTA / FASTQ reading (gzip-compressed or plain) and the
// GFA writer.
#pragma once

#include <ostream>
#include <string>
#include <vector>

#include "spoa/graph.hpp"

namespace spoa_amd {

struct SeqRecord {
  std::string name;
  std::string data;
  std::string quality;  // empty for FASTA
};

// True when the path has one of the extensions the reference's tool accepts
// (src/main.cpp:25-59); *fastq says which format.
bool SequenceFormat(const std::string& path, bool* fastq);

// Every record of a FASTA or FASTQ file (gzip or plain).  A record's name is
// its header line after the '>' / '@', trailing whitespace removed; FASTA
// sequences may span several lines.  Throws std::invalid_argument on a file
// it cannot read or parse.
std::vector<SeqRecord> ReadSequences(const std::string& path);

// The tool's GFA output (src/main.cpp:119-199): a segment per node (the
// consensus nodes tagged), a link per edge with its weight, a path per
// sequence (reversed ones walked backwards with '-' orientations), and with
// include_consensus a path of the consensus.  Uses the graph's last
// consensus.  Returns false (nothing written) when headers or flags are
// missing.
bool WriteGfa(std::ostream& os, const spoa::Graph& graph, const std::vector<std::string>& headers,
              const std::vector<bool>& is_reversed, bool include_consensus);

}  // namespace spoa_amd
