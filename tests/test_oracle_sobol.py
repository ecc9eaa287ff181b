"""Sobol oracle (oracle/sobol.c) against the reference's known-answer tests
(shared-samplers/src/__tests__/sobol.test.ts:57-143), plus stratification of
the unscrambled sequence."""
import numpy as np

from oracle import oracle as O

S = O.sobol


def test_bit_reversal_and_hash_kats():
    assert S("rev32", 0x00000001) == 0x80000000
    assert S("rev32", 0x80000000) == 0x00000001
    assert S("rev32", 0x0F0F0F0F) == 0xF0F0F0F0
    assert S("sobol_hash", 1) == 0x514E28B7
    assert S("sobol_combine", 0x12345678, 0x9ABCDEF0) == 0x3E6BC4F6
    assert S("laine_karras", 0x13579BDF, 0x2468ACE0) == 0x0A4633FB
    assert S("nested_scramble", 0x00ABCDEF, 0x31415926) == 0x81C5A4EA


def test_rotation_tile_kats():
    assert S("sobol_rotation", 0, 9) == 0
    assert S("sobol_rotation", 1, 9) == 0x89A213
    assert S("sobol_rotation", 4, 9) == 0xBCDA23
    assert S("sobol_rotation", 63, 9) == 0xE40131
    assert S("sobol_rotation", 0, 323) == 0


def test_direction_table_kats():
    assert S("sobol_component", 7, 0) == 7
    assert S("sobol_component", 7, 5) == 7
    assert S("sobol_component", 7, 255) == 6
    assert S("sobol_component", 7, 511) == 6
    pts = {1: [1, 1, 1, 1], 2: [3, 3, 3, 2], 3: [2, 2, 2, 3], 4: [5, 6, 4, 4]}
    for i, comps in pts.items():
        assert [S("sobol_component", i, d) for d in range(4)] == comps
    assert [S("sobol_component", 0, d) for d in range(4)] == [0, 0, 0, 0]


def test_owen_scrambled_kats():
    assert S("sobol_owen_u32", 0, 0, 0, 0) == 0xA66DE000
    assert S("sobol_owen_u32", 0, 1, 0, 0) == 0x1FC82700
    assert S("sobol_owen_u32", 1, 0, 0, 0) == 0x5336F000
    v = S("sobol_owen_u32", 12345, 3, 63, 0)
    assert v == 0xBA554D00
    assert abs((v >> 8) * 2.0 ** -24 - 0.7278640866279602) < 1e-12


def test_stream_kats():
    key = S("sobol_frame_key", 123, 0)
    assert key == 0xC00C0000
    assert O.sobol_stream_state(9, 10, key) == {
        "sampleIndex": 0, "dimension": 0, "pixelX": 9, "pixelY": 10, "sequenceKey": 49164,
        "rotationTile": 17, "fallbackState": 3105986336}
    words = O.sobol_streams([9], [10], key, 8)[0]
    assert [int(w) for w in words] == [0x8AECA400, 0x419E9400, 0xB524E900, 0x23E98800, 0x7759A400,
                                       0xE7614000, 0xF4738A00, 0xB225F000]


def test_pixels_do_not_alias_and_dimensions_differ():
    key = S("sobol_frame_key", 0x12345678, 99)
    a, b = O.sobol_streams([0, 8], [0, 0], key, 16)
    assert not np.array_equal(a, b)
    assert S("sobol_owen_u32", 12345, 0, 0, 0) != S("sobol_owen_u32", 12345, 3, 0, 0)


def test_unscrambled_components_are_stratified():
    """Components are stored bit-reversed (16 bits); reversed back, each of
    the first 64 dimensions puts exactly one of the first 2^k points in each
    of the 2^k dyadic intervals (a (0,k,1)-net per dimension)."""
    for d in range(64):
        for k in (4, 8, 12):
            vals = np.array([S("rev32", S("sobol_component", i, d)) >> 16 for i in range(1 << k)], np.uint64)
            cells = vals >> np.uint64(16 - k)
            assert np.array_equal(np.sort(cells), np.arange(1 << k)), (d, k)
