"""Oracle pins: RNG and transcendentals.

- PCG32 against the published pcg32-demo output for (initstate 42, initseq 54)
  and against an independent Python big-integer statement of O'Neill's
  algorithm (multiplier 6364136223846793005, XSH-RR output).
- PCG-hash (Jarzynski & Olano 2020) is a bijection on 32-bit words: a wrong
  constant or shift that breaks invertibility shows up as collisions.
- The portable sin/cos/log/exp/pow stay within a few ulp of float64 numpy on
  the domains the integrator uses.
"""
import numpy as np

from oracle import pyoracle as O

M64 = (1 << 64) - 1


def pcg32_ref(initstate, initseq, n):
    state = 0
    inc = ((initseq << 1) | 1) & M64

    def step(s):
        return (s * 6364136223846793005 + inc) & M64

    def out(old):
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    state = step(state)
    state = (state + initstate) & M64
    state = step(state)
    res = []
    for _ in range(n):
        old = state
        state = step(state)
        res.append(out(old))
    return res


def test_pcg32_published_vector():
    got = O.pcg32(42, 54, 6)
    assert [int(x) for x in got] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_pcg32_bigint_restatement():
    for st, sq in [(0, 0), (1, 2), (0xDEADBEEF, 77), (2**63 + 5, 2**40 + 3)]:
        assert [int(x) for x in O.pcg32(st, sq, 500)] == pcg32_ref(st, sq, 500)


def test_path_floats_are_top_24_bits():
    pix, seed, sample = 1234, 7, 3
    f = O.path_floats(pix, seed, sample, 64)
    raw = pcg32_ref(O.seed_hash(pix, seed), sample, 64)
    expect = np.array([(r >> 8) * 2.0**-24 for r in raw], np.float32)
    assert np.array_equal(f, expect)
    assert f.min() >= 0.0 and f.max() < 1.0


def test_seed_hash_bijective_and_mixing():
    vals = np.array([O.seed_hash(p, 0) for p in range(1 << 15)], np.uint64)
    assert np.unique(vals).size == vals.size
    # adjacent pixels differ in about half the bits
    x = vals[:-1] ^ vals[1:]
    bits = np.unpackbits(x.astype(">u4").view(np.uint8)).reshape(-1, 32).sum(1)
    assert 14 < bits.mean() < 18


def _ulp_err(got, ref):
    ref32 = ref.astype(np.float32)
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / spacing


def test_sincos_ulp():
    x = np.linspace(-np.pi, 2 * np.pi, 200001).astype(np.float32)
    for fn, ref in (("sin", np.sin), ("cos", np.cos)):
        got = O.mathf(fn, x)
        r = ref(x.astype(np.float64))
        big = np.abs(r) > 1e-3
        assert _ulp_err(got[big], r[big]).max() <= 4, fn
        assert np.abs(got - r).max() < 2e-7, fn


def test_log_exp_pow_ulp():
    x = np.exp(np.linspace(-60, 60, 100001)).astype(np.float32)
    got = O.mathf("log", x)
    r = np.log(x.astype(np.float64))
    big = np.abs(r) > 1e-2
    assert _ulp_err(got[big], r[big]).max() <= 4
    small = np.abs(r) < 1.0
    assert np.abs(got[small] - r[small]).max() < 1.5e-7
    e = np.linspace(-80, 80, 100001).astype(np.float32)
    assert _ulp_err(O.mathf("exp", e), np.exp(e.astype(np.float64))).max() <= 4
    b = np.linspace(1e-4, 4.0, 50001).astype(np.float32)
    for y in (1.0 / 2.2, 0.5, 0.25, 0.999):
        y32 = np.float32(y)
        got = O.mathf("pow", b, y32)
        ref = np.power(b.astype(np.float64), np.float64(y32))
        # exp(y log x): the rounding of y log x is amplified by |y log x|
        bound = 4.0 + np.abs(np.float64(y32) * np.log(b.astype(np.float64)))
        assert np.all(_ulp_err(got, ref) <= bound), y


def test_math_special_values():
    assert np.isneginf(O.mathf("log", np.array([0.0], np.float32)))[0]
    assert np.isnan(O.mathf("log", np.array([-1.0], np.float32)))[0]
    assert O.mathf("exp", np.array([0.0], np.float32))[0] == 1.0
    assert O.mathf("pow", np.array([0.0], np.float32), np.float32(0.45))[0] == 0.0
    assert np.isnan(O.mathf("pow", np.array([-0.5], np.float32), np.float32(0.45)))[0]
    assert O.mathf("pow", np.array([1.0], np.float32), np.float32(0.45))[0] == 1.0
