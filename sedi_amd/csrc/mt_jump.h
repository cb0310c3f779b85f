// mt_jump.h -- mt19937 jump-ahead (host).
#pragma once

#include <cstdint>
#include <vector>

namespace sedi {

// The 624-word state of std::mt19937(seed) before its first output.
void mt_seed_state(uint32_t seed, uint32_t* s);
// x^J mod phi (cached per J).
const std::vector<uint64_t>& mt_jump_poly(unsigned long long J);
// out = the state J outputs after s (both oldest word first; out may not alias s).
void mt_jump_state(const uint32_t* s, unsigned long long J, uint32_t* out);
bool mt_charpoly_ok();
// x^(2*ncells*k) mod phi: the jump from a stream's start to probe k of a grid
// of ncells cells (two outputs per cell).  Cached; independent of the seed.
const std::vector<uint64_t>& mt_probe_poly(long long ncells, int k);
// x^(J*j) mod phi for j = 0 .. count-1, concatenated (count * mt_poly_words()
// words; cached per (J, count)).
void mt_poly_powers(unsigned long long J, int count, std::vector<uint64_t>& out);
// Words per cached polynomial (19937 bits rounded up, plus one).
int mt_poly_words();
// Untempered word sequence w_0 .. w_{count-1} from state s (w_0..w_623 = s).
void mt_word_sequence(const uint32_t* s, uint32_t* w, int count);
// Degree of the characteristic polynomial and the words of one jump window.
constexpr int MT_DEG = 19937;
constexpr int MT_WINDOW = MT_DEG + 624;

}  // namespace sedi
