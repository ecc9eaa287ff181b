// sobol.cpp -- Sobol direction components from the Joe-Kuo D(6) rows.
#include "sobol.h"

#include <cstring>

namespace {

struct JoeKuoRow {
  uint32_t src, degree, poly, m[13];
};

const JoeKuoRow kRows[vts::kDims] = {
#include "joe_kuo_d6.inc"
};

}  // namespace

namespace vts {

// Direction numbers by the Joe-Kuo recurrence (sobolDirectionNumbers.ts:
// 578-611): v_k = m_k 2^(31-k) for k < s, then
// v_k = v_{k-s} ^ (v_{k-s} >> s) ^ sum_{j<s, a_j} v_{k-j}; stored bit-reversed,
// low 16 bits.
void build_directions(uint16_t* out) {
  for (uint32_t d = 0; d < kDims; ++d) {
    const JoeKuoRow& r = kRows[d];
    uint32_t v[kBits];
    std::memset(v, 0, sizeof v);
    for (uint32_t b = 0; b < r.degree && b < kBits; ++b) v[b] = r.m[b] << (31u - b);
    for (uint32_t b = r.degree; b < kBits; ++b) {
      uint32_t x = v[b - r.degree] ^ (v[b - r.degree] >> r.degree);
      for (uint32_t k = 1; k < r.degree; ++k)
        if ((r.poly >> (r.degree - 1u - k)) & 1u) x ^= v[b - k];
      v[b] = x;
    }
    for (uint32_t b = 0; b < kBits; ++b) out[d * kBits + b] = (uint16_t)(rev32(v[b]) & 0xFFFFu);
  }
}

}  // namespace vts
