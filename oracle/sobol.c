/*
 * sobol.c -- Owen-scrambled Sobol streams of pt-webgpu's binding-free Sobol
 * sampler (packages/shared-samplers/src/sobol.ts:55-245,
 * sobolDirectionNumbers.ts:568-640).  TEST INFRASTRUCTURE ONLY (see
 * vt_oracle.h).
 *
 * Joe-Kuo D(6) direction numbers (dimensions 2..513 of the source file map to
 * sampler dimensions 0..511), 16 direction bits, Laine-Karras nested uniform
 * scrambling, an additive 8x8 ranked rotation tile and a PCG continuation
 * after the last dimension.
 */
#include <stdint.h>
#include <string.h>

typedef struct {
  uint32_t src, degree, poly, m[13];
} jk_row;

static const jk_row JK_ROWS[512] = {
#include "../vitrum_amd/csrc/joe_kuo_d6.inc"
};

/* sobol.ts:30-39 */
static const uint8_t RANK8X8[64] = {
    0,  63, 12, 60, 3,  55, 15, 62, 53, 23, 57, 17, 44, 19, 32, 22, 10, 40, 5,  41, 8,  35,
    7,  47, 45, 28, 48, 25, 54, 29, 36, 24, 2,  38, 13, 46, 1,  37, 14, 51, 58, 30, 49, 16,
    59, 20, 43, 18, 11, 56, 6,  34, 9,  39, 4,  50, 52, 31, 33, 27, 42, 26, 61, 21};

static uint16_t g_comp[512 * 16];
static int g_ready = 0;

uint32_t orc_rev32(uint32_t x) {
  x = ((x & 0xAAAAAAAAu) >> 1) | ((x & 0x55555555u) << 1);
  x = ((x & 0xCCCCCCCCu) >> 2) | ((x & 0x33333333u) << 2);
  x = ((x & 0xF0F0F0F0u) >> 4) | ((x & 0x0F0F0F0Fu) << 4);
  x = ((x & 0xFF00FF00u) >> 8) | ((x & 0x00FF00FFu) << 8);
  return (x >> 16) | (x << 16);
}

/* sobolDirectionNumbers.ts:578-611: direction numbers, stored bit-reversed
 * and truncated to the low 16 bits (the component domain). */
static void build_table(void) {
  if (g_ready) return;
  for (uint32_t d = 0; d < 512; ++d) {
    const jk_row *r = &JK_ROWS[d];
    uint32_t v[16];
    memset(v, 0, sizeof v);
    for (uint32_t b = 0; b < r->degree && b < 16; ++b) v[b] = r->m[b] << (31u - b);
    for (uint32_t b = r->degree; b < 16; ++b) {
      uint32_t x = v[b - r->degree] ^ (v[b - r->degree] >> r->degree);
      for (uint32_t k = 1; k < r->degree; ++k)
        if ((r->poly >> (r->degree - 1u - k)) & 1u) x ^= v[b - k];
      v[b] = x;
    }
    for (uint32_t b = 0; b < 16; ++b) g_comp[d * 16 + b] = (uint16_t)(orc_rev32(v[b]) & 0xFFFFu);
  }
  g_ready = 1;
}

uint32_t orc_sobol_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}

uint32_t orc_sobol_combine(uint32_t s, uint32_t v) { return s ^ (v + ((s << 6) + (s >> 2))); }

uint32_t orc_laine_karras(uint32_t v, uint32_t seed) {
  uint32_t x = v + seed;
  x ^= x * 0x6C50B47Cu;
  x ^= x * 0xB82F1E52u;
  x ^= x * 0xC7AFE638u;
  x ^= x * 0x8D22F6E6u;
  return x;
}

uint32_t orc_nested_scramble(uint32_t v, uint32_t seed) { return orc_rev32(orc_laine_karras(v, seed)); }

/* sobolTextureComponentBits: 24-bit component of point `index` (mod 2^16). */
uint32_t orc_sobol_component(uint32_t index, uint32_t dim) {
  build_table();
  if (dim >= 512) return 0;
  const uint32_t i = index & 0xFFFFu;
  uint32_t out = 0;
  for (uint32_t b = 0; b < 16; ++b)
    if ((i >> b) & 1u) out ^= g_comp[dim * 16 + b];
  return out & 0x00FFFFFFu;
}

uint32_t orc_sobol_rotation(uint32_t tile, uint32_t dim) {
  const uint32_t rank = RANK8X8[tile & 63u];
  if (rank == 0) return 0;
  return orc_sobol_hash(orc_sobol_combine(rank, dim)) & 0x00FFFFFFu;
}

uint32_t orc_sobol_frame_key(uint32_t seed, uint32_t frame) {
  const uint32_t block = frame >> 16;
  const uint32_t seed_key = orc_sobol_hash(seed) & 0xFFFFu;
  const uint32_t block_key = (seed_key + block * 0x9E37u) & 0xFFFFu;
  return (block_key << 16) | (frame & 0xFFFFu);
}

/* owenScrambledSobolU32 (sobol.ts:187-210) */
uint32_t orc_sobol_owen_u32(uint32_t path, uint32_t dim, uint32_t tile, uint32_t stream_seed) {
  path &= 0xFFFFu;
  const uint32_t seed = orc_sobol_hash(orc_sobol_combine(stream_seed, dim));
  const uint32_t shuffle_seed = orc_sobol_combine(stream_seed, 0);
  const uint32_t shuffled = orc_nested_scramble(orc_rev32(path), shuffle_seed) & 0xFFFFu;
  const uint32_t result = orc_sobol_component(shuffled, dim);
  const uint32_t comp_seed = orc_sobol_combine(seed, 1u + (dim & 3u));
  const uint32_t s24 = (orc_nested_scramble(result, comp_seed) >> 8) & 0x00FFFFFFu;
  const uint32_t r24 = (s24 + orc_sobol_rotation(tile, dim)) & 0x00FFFFFFu;
  return r24 << 8;
}

typedef struct {
  uint32_t sample_index, dimension, px, py, sequence_key, rotation_tile, fallback;
} orc_sobol_stream;

/* initOwenScrambledSobolStream (sobol.ts:132-157) */
void orc_sobol_init(orc_sobol_stream *s, uint32_t px, uint32_t py, uint32_t key) {
  s->sample_index = key & 0xFFFFu;
  s->dimension = 0;
  s->px = px;
  s->py = py;
  s->sequence_key = key >> 16;
  s->rotation_tile = (px & 7u) | ((py & 7u) << 3);
  const uint32_t pixel_seed = orc_sobol_hash(orc_sobol_combine(orc_sobol_hash(px), py));
  s->fallback = orc_sobol_hash(orc_sobol_combine(
      orc_sobol_combine(orc_sobol_combine(pixel_seed, s->sequence_key), s->sample_index), 512u));
}

/* nextOwenScrambledSobolU32 (sobol.ts:220-238), PCG continuation past 512. */
uint32_t orc_sobol_next(orc_sobol_stream *s) {
  const uint32_t dim = s->dimension;
  if (dim >= 512u) {
    if (dim != 0xFFFFFFFFu) s->dimension = dim + 1u;
    const uint32_t next = s->fallback * 747796405u + 2891336453u;
    s->fallback = next;
    uint32_t w = ((next >> ((next >> 28) + 4u)) ^ next) * 277803737u;
    return (w >> 22) ^ w;
  }
  const uint32_t pixel_seed = orc_sobol_hash(orc_sobol_combine(orc_sobol_hash(s->px), s->py));
  const uint32_t stream_seed = orc_sobol_hash(orc_sobol_combine(pixel_seed, s->sequence_key));
  const uint32_t v = orc_sobol_owen_u32(s->sample_index, dim, s->rotation_tile, stream_seed);
  s->dimension = dim + 1u;
  return v;
}

/* Stream words: out[i * draws + k] = k-th word of pixel (px[i], py[i]). */
void orc_sobol_streams(const uint32_t *px, const uint32_t *py, uint32_t n, uint32_t key, uint32_t draws,
                       uint32_t *out) {
  for (uint32_t i = 0; i < n; ++i) {
    orc_sobol_stream s;
    orc_sobol_init(&s, px[i], py[i], key);
    for (uint32_t k = 0; k < draws; ++k) out[(size_t)i * draws + k] = orc_sobol_next(&s);
  }
}

void orc_sobol_stream_state(uint32_t px, uint32_t py, uint32_t key, uint32_t *state7) {
  orc_sobol_stream s;
  orc_sobol_init(&s, px, py, key);
  memcpy(state7, &s, sizeof s);
}
