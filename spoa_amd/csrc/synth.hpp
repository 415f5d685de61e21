// Seeded synthetic POA windows (template + noisy reads).
//
// The same generator feeds the GPU path, the CPU baselines and the test
// fixtures, so every consumer sees identical bytes for a given seed.
//
// PRNG: splitmix64 expands the 64-bit seed into the state of xoshiro256**
// (Blackman & Vigna, "Scrambled linear pseudorandom number generators",
// 2018).  Integer ranges use the high 32 bits times the range (no
// rejection); reals use the top 53 bits.
//
// Error model (per source base, probability p): substitution to one of the
// three other bases, insertion of a uniform random base before the source
// base, or deletion, each with probability 1/3.  Read sources are either
// the whole template or a substring of it at a uniform offset, of the drawn
// length clipped to the template length.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace spoa_synth {

struct Rng {
  std::uint64_t s[4];

  static std::uint64_t SplitMix(std::uint64_t* x) {
    std::uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  explicit Rng(std::uint64_t seed) {
    for (auto& v : s) v = SplitMix(&seed);
  }
  static std::uint64_t Rotl(std::uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  std::uint64_t Next() {
    const std::uint64_t result = Rotl(s[1] * 5, 7) * 9;
    const std::uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = Rotl(s[3], 45);
    return result;
  }
  // uniform integer in [lo, hi], hi - lo < 2^32
  std::uint32_t Int(std::uint32_t lo, std::uint32_t hi) {
    std::uint64_t span = static_cast<std::uint64_t>(hi) - lo + 1;
    return lo + static_cast<std::uint32_t>(((Next() >> 32) * span) >> 32);
  }
  double Real() { return static_cast<double>(Next() >> 11) * (1.0 / 9007199254740992.0); }
  char Base() { return "ACGT"[Next() >> 62]; }
};

struct WindowSpec {
  std::uint32_t template_min, template_max;  // template length range
  std::uint32_t reads_min, reads_max;        // reads per window
  std::uint32_t read_len_min, read_len_max;  // 0,0: reads span the whole template
  double err_min, err_max;                   // per-window error rate range
  std::uint32_t revcomp_odd;                 // 1: odd-numbered reads are reverse-complemented
  std::uint32_t log_uniform_len;             // 1: read length log-uniform in its range (the template's
                                             // when reads span the template)
};

struct Window {
  std::string templ;
  std::vector<std::string> reads;
  std::vector<std::uint8_t> reversed;
  double err = 0;
};

inline char Complement(char b) {
  switch (b) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return b;
  }
}

inline void ReverseComplement(std::string* s) {
  std::string r(s->rbegin(), s->rend());
  for (auto& b : r) b = Complement(b);
  s->swap(r);
}

inline std::string Mutate(Rng* rng, const std::string& src, double p) {
  std::string out;
  out.reserve(src.size() + src.size() / 8 + 8);
  for (char b : src) {
    if (rng->Real() < p) {
      std::uint32_t kind = rng->Int(0, 2);
      if (kind == 0) {
        char nb;
        do { nb = rng->Base(); } while (nb == b);
        out.push_back(nb);
      } else if (kind == 1) {
        out.push_back(rng->Base());
        out.push_back(b);
      }  // kind 2: deletion
    } else {
      out.push_back(b);
    }
  }
  return out;
}

// Seed of window w of a seeded set: one splitmix64 step over seed and w.
inline std::uint64_t WindowSeed(std::uint64_t seed, std::uint64_t w) {
  std::uint64_t x = seed ^ (0xD1B54A32D192ED03ULL * (w + 1));
  return Rng::SplitMix(&x);
}

inline Window MakeWindow(std::uint64_t seed, const WindowSpec& sp) {
  Rng rng(seed);
  Window w;
  std::uint32_t tlen;
  if (sp.log_uniform_len && sp.read_len_max == 0) {
    // reads span the template: the template length is the log-uniform one
    const double lo = static_cast<double>(sp.template_min), hi = static_cast<double>(sp.template_max);
    tlen = static_cast<std::uint32_t>(lo * std::exp(rng.Real() * std::log(hi / lo)) + 0.5);
  } else {
    tlen = rng.Int(sp.template_min, sp.template_max);
  }
  w.templ.resize(tlen);
  for (auto& b : w.templ) b = rng.Base();
  std::uint32_t nreads = rng.Int(sp.reads_min, sp.reads_max);
  w.err = sp.err_min + (sp.err_max - sp.err_min) * rng.Real();
  for (std::uint32_t r = 0; r < nreads; ++r) {
    std::string src;
    if (sp.read_len_max == 0) {
      src = w.templ;
    } else {
      std::uint32_t len;
      if (sp.log_uniform_len) {
        double lo = static_cast<double>(sp.read_len_min), hi = static_cast<double>(sp.read_len_max);
        double u = rng.Real();
        len = static_cast<std::uint32_t>(lo * std::exp(u * std::log(hi / lo)) + 0.5);
      } else {
        len = rng.Int(sp.read_len_min, sp.read_len_max);
      }
      // a substring of the template; reads longer than the template are
      // the whole template
      len = std::min(len, tlen);
      const std::uint32_t off = rng.Int(0, tlen - len);
      src = w.templ.substr(off, len);
    }
    std::string read = Mutate(&rng, src, w.err);
    if (read.empty()) read.push_back(rng.Base());
    std::uint8_t rev = 0;
    if (sp.revcomp_odd && (r & 1)) {
      ReverseComplement(&read);
      rev = 1;
    }
    w.reads.push_back(std::move(read));
    w.reversed.push_back(rev);
  }
  return w;
}

}  // namespace spoa_synth
