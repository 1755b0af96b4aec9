// Variants file: the overrides of a mutant / perturbation sweep, one block per variant.
//
//   // comment
//   variant NAME {
//       NODE.knockout;                 istate 0, rate_up x 0: the node stays OFF
//       NODE.overexpress;              istate 1, rate_down x 0: the node stays ON
//       NODE.rate_up *= NUMBER;        multiplies the node's rate_up (repeated lines compose)
//       NODE.rate_down *= NUMBER;
//       NODE.istate = NUMBER;          probability that the node is ON at t = 0, in [0, 1]
//       sample_count = INTEGER;        trajectories of this variant (default: the model's)
//       seed_pseudorandom = INTEGER;   seed of this variant (default: the model's)
//   }
//
// NAME is [A-Za-z_][A-Za-z0-9_-]* and unique in the file.  An empty block is the model
// itself.  The reference has no counterpart (its README.md:33 lists mutant analysis as future
// work); the knockout / overexpression semantics are MaBoSS's usual mutant forms (rate_up or
// rate_down forced to 0 and the initial state fixed accordingly).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mbg {

struct VariantSpec {
    std::string name;
    int line = 0;                              // line of the block header
    std::vector<float> mult_up, mult_down;     // [N] factors on the base model's multipliers
    std::vector<float> istate;                 // [N] override, or NaN: the base model's value
    bool has_count = false, has_seed = false;
    uint64_t count = 0, seed = 0;
};

// Parses a variants file for a model whose nodes are `names` (in .bnd order).  Throws
// ParseError with "line L: ..." on any error.
std::vector<VariantSpec> parse_variants(const std::string& text, const std::vector<std::string>& names);

}  // namespace mbg
