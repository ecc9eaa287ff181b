"""PCG sampler of the oracle (pcgInit / pcgNext / rand_f32,
shared-samplers/src/wgsl/pcg.wgsl.ts:5-22) pinned without restating it.

The reference ships no PCG known answers, so each output word is run
backwards: the RXS-M-XS output step is a bijection on 32-bit words (two
xorshifts and a multiplication by an odd constant), and undoing it recovers
the generator's state.  Those states must then follow the published PCG
32-bit LCG (O'Neill 2014, pcg-c's default multiplier 747796405 and
increment 2891336453) step by step and by the closed-form jump
a^k s + c (a^k - 1) / (a - 1); and pcgInit's state, with its three
xorshifts undone, must equal the linear hash px*1664525 + py*1013904223 +
key*22695477 (the Numerical Recipes LCG constants).  A wrong shift, constant
or step order in the generator fails one of these; the inverses share no
code with it.
"""
import numpy as np

from oracle import oracle as O
from vitrum_amd import workloads as W

M32 = 0xFFFFFFFF
LCG_A, LCG_C = 747796405, 2891336453
RXS_MUL = 277803737


def undo_xorshift_right(y, k):
    x = y
    for _ in range(32 // k + 1):
        x = y ^ (x >> k)
    return x


def undo_xorshift_left(y, k):
    x = y
    for _ in range(32 // k + 1):
        x = y ^ ((x << k) & M32)
    return x


def state_of_word(w):
    """Undo the RXS-M-XS output step: the state the word was drawn from."""
    x = w ^ (w >> 22)
    m = (x * pow(RXS_MUL, -1, 1 << 32)) & M32
    return undo_xorshift_right(m, (m >> 28) + 4)


def lcg_jump(s, k):
    """a^k s + c (a^k - 1) / (a - 1) mod 2^32, the geometric sum taken modulo
    2^32 (a - 1) so that the division is exact."""
    mod = (1 << 32) * (LCG_A - 1)
    ak = pow(LCG_A, k, mod)
    return (ak * s + LCG_C * ((ak - 1) // (LCG_A - 1))) & M32


TRIPLES = [(0, 0, 0), (9, 10, 0x9E3779B9), (1023, 511, 7), (4096, 3, 0xDEADBEEF), (M32, M32, M32)]


def test_init_is_the_linear_hash_with_xorshifts():
    for px, py, key in TRIPLES:
        s = O.pcg_init(px, py, key)
        t = undo_xorshift_right(undo_xorshift_left(undo_xorshift_right(s, 11), 31), 17)
        assert t == (px * 1664525 + py * 1013904223 + key * 22695477) & M32
        assert int(W.pcg_init(px, py, key)) == s  # the product's ray generator


def test_words_invert_to_consecutive_lcg_states():
    for px, py, key in TRIPLES:
        s0 = O.pcg_init(px, py, key)
        words = O.pcg_stream(px, py, key, 64)
        states = [state_of_word(w) for w in words]
        assert states[0] == (LCG_A * s0 + LCG_C) & M32
        for a, b in zip(states, states[1:]):
            assert b == (LCG_A * a + LCG_C) & M32
        assert states[-1] == lcg_jump(s0, 64)


def test_long_stream_follows_the_jump_formula():
    words = O.pcg_stream(3, 5, 11, 5000)
    s0 = O.pcg_init(3, 5, 11)
    for k in (1, 2, 17, 1000, 5000):
        assert state_of_word(words[k - 1]) == lcg_jump(s0, k)
    # a 5000-word stream of a bijective output of a full-period LCG has no
    # repeated states, and its top bits are balanced
    assert len(set(words)) == 5000
    top = np.array(words, np.uint64) >> np.uint64(31)
    assert abs(int(top.sum()) - 2500) < 4 * np.sqrt(1250)


def test_output_step_is_a_bijection():
    rng = np.random.default_rng(3)
    for s in rng.integers(0, 1 << 32, 2000, dtype=np.uint64).tolist():
        st = np.uint32(s)
        _, w = W.pcg_next(np.uint32((int(st) - LCG_C) * pow(LCG_A, -1, 1 << 32) & M32))
        assert state_of_word(int(w)) == s


def test_rand_f32_is_the_top_24_bits():
    s = np.uint32(O.pcg_init(9, 10, 0x9E3779B9))
    words = O.pcg_stream(9, 10, 0x9E3779B9, 8)
    for w in words:
        s, f = W.rand_f32(s)
        assert float(f) == (w >> 8) / 16777216.0
