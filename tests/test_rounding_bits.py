"""Rounding bits of the asynchronous kernel's W-row stores (bf16, stochastic).

nomad_amd/csrc/async_kernel.hip draws them as lanekey + acc, where acc is a
per-wave scalar advanced by 0x9E3779B9 before every rating and started from
mix32(seed ^ mix32(wave * 0x9E3779B1 + launch)); lanekey =
mix32(seed ^ (lane + 1) * 0x85EBCA6B), the same key as the hashed bits of
the other kernels.  Element 2l takes the low 16 bits of the sum, element
2l+1 the high 16 bits.  Unlike the hashed bits (bit-exact against the oracle
in strat mode), these are only compared through the RMSE of the async
tests, so what is checked here is the property stochastic rounding needs:
for every lane, each half of the bits makes a value round up with
probability equal to its position inside the bf16 ulp, and the two halves
of a word round independently.  A plausible slip (the scalar not advanced
per rating, a step whose low 16 bits are zero, or both elements taking the
same half) fails one of the checks below.
"""
import numpy as np

M32 = 0xFFFFFFFF
PHI = 0x9E3779B9


def mix32(x):
    x = np.asarray(x, dtype=np.uint64) & M32
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def weyl_bits(seed, wave, launch, lane, n):
    """The n rounding words of `lane` for the first n ratings of a wave."""
    acc0 = int(mix32(seed ^ int(mix32((wave * 0x9E3779B1 + launch) & M32))))
    key = int(mix32(seed ^ (((lane + 1) * 0x85EBCA6B) & M32)))
    steps = (np.arange(1, n + 1, dtype=np.uint64) * PHI + acc0) & M32
    return (steps + key) & M32


def round_up(x_low16, r16):
    """Does f32_to_bf16_sr carry into the kept bits (round up)?"""
    return (x_low16 + r16) >= 0x10000


def test_weyl_rounding_bits_unbiased_per_lane():
    n = 20000
    seed = 0xA5A5A5A5 ^ 12345
    for frac in (0.1, 0.3, 0.5, 0.77):
        x_low16 = int(frac * 0x10000)
        p = x_low16 / 0x10000
        sigma = np.sqrt(p * (1 - p) / n)
        for wave, launch, lane in ((0, 0, 0), (7, 3, 31), (2047, 11, 63), (123, 0, 5)):
            r = weyl_bits(seed, wave, launch, lane, n)
            for half in (r & 0xFFFF, r >> 16):
                p_hat = round_up(x_low16, half).mean()
                assert abs(p_hat - p) < 4 * sigma, (frac, wave, launch, lane, p_hat)


def test_weyl_halves_of_a_word_independent():
    """The two elements of a pair do not round alike: over many ratings
    their round-up decisions agree about as often as independent draws
    would (one half for p = 0.5).  (Lanes are another matter: like the
    hashed bits, whose lanes differ by a fixed XOR key, the lanes here
    differ by a fixed added key, so a given pair of lanes rounds in a fixed
    relation; and consecutive ratings of one lane, which update different
    rows, step by the golden ratio, so their decisions alternate more than
    independent draws would.  Neither biases a rounding.)"""
    n = 20000
    x = 0x8000
    for seed, wave, launch, lane in ((0x12345678, 5, 1, 10), (3, 900, 2, 63), (99, 0, 0, 0)):
        a = weyl_bits(seed, wave, launch, lane, n)
        agree = (round_up(x, a & 0xFFFF) == round_up(x, a >> 16)).mean()
        assert abs(agree - 0.5) < 0.03, (seed, wave, launch, lane, agree)


def test_weyl_sequence_distinct_per_wave_and_launch():
    seed = 7
    w0 = weyl_bits(seed, 0, 0, 0, 64)
    assert not np.array_equal(w0, weyl_bits(seed, 1, 0, 0, 64))
    assert not np.array_equal(w0, weyl_bits(seed, 0, 1, 0, 64))
    assert len(np.unique(w0)) == 64
