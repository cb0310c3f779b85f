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

}  // namespace sedi
