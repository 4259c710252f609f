"""Pins of the oracle's random numbers and spectral arithmetic against published values and
closed forms (no GPU).

  PCG32RNG (random_distributions.h:6-27)         pcg32 demo vectors, LCG closed form,
                                                  getFloat0cTo1o = top 23 bits / 2^23
  per-pixel seeding (context.cpp:1182-1187)       C++ standard mt19937_64 check value
  CIE 1931 table (spectrum_base.cpp:6-190)        published CIE values and sums
  CMF strata (spectrum_types.cpp:624-670)         integral of the piecewise-linear table
  Meng upsampling (spectrum_types.cpp:11-91, 273-336)
                                                  gray flatness; XYZ of the upsampled
                                                  spectrum = Rec.709 primaries / white point
                                                  transform of the input (matrices derived here
                                                  from the published chromaticities)
"""
import json
import os

import numpy as np
import pytest

from oracle import binding as ob

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "published_kats.json")))
MASK = (1 << 64) - 1
MULT = 6364136223846793005
TABLES = os.path.join(os.path.dirname(HERE), "vlr_amd", "data", "spectral_tables.bin")


def cmf_table():
    b = open(TABLES, "rb").read()
    assert b[:8] == b"VLRSPEC1"
    n = int(np.frombuffer(b[8:12], np.uint32)[0])
    return np.frombuffer(b[12:12 + 12 * n], np.float32).reshape(3, n).astype(np.float64)


def rgb_to_xyz_matrix(primaries, white):
    """Columns = XYZ of the primaries scaled so that RGB (1,1,1) maps to the white (Y = 1)."""
    P = np.array([[x / y, 1.0, (1 - x - y) / y] for x, y in primaries]).T
    W = np.array([white[0] / white[1], 1.0, (1 - white[0] - white[1]) / white[1]])
    return P * np.linalg.solve(P, W)


# ---------------------------------------------------------------------------------------------
def test_pcg32_output_function_matches_published_demo():
    k = KATS["pcg32_demo"]
    inc = (k["initseq"] << 1) | 1
    state = (inc) & MASK                       # srandom: state = 0, step
    state = (state + k["initstate"]) & MASK
    state = (state * MULT + inc) & MASK
    for expected in k["outputs"]:
        # the output of a step depends only on the state before it
        out, _ = ob.pcg32_stream(state, 1)
        assert int(out[0]) == expected
        state = (state * MULT + inc) & MASK


@pytest.mark.parametrize("seed", [0, 1, 591842031321323413, MASK])
def test_pcg32_state_closed_form(seed):
    # increment 1: s_n = M^n s_0 + (M^n - 1) / (M - 1), evaluated exactly as sum_{k<n} M^k
    n = 100_000
    _, final = ob.pcg32_stream(seed, n)
    mn = pow(MULT, n, 1 << 64)
    # geometric sum mod 2^64 via (M^n - 1) / (M - 1) over integers modulo 2^64 * (M - 1)
    geo = ((pow(MULT, n, (1 << 64) * (MULT - 1)) - 1) // (MULT - 1)) & MASK
    assert final == (mn * seed + geo) & MASK


def test_pcg32_float_is_top_23_bits():
    for seed in (7, 12345, 2 ** 63 + 11):
        ints, _ = ob.pcg32_stream(seed, 4096)
        floats = ob.pcg32_floats(seed, 4096)
        expect = (ints >> 9).astype(np.float64) / 2.0 ** 23
        assert np.array_equal(floats.astype(np.float64), expect)
        assert floats.min() >= 0.0 and floats.max() < 1.0


def test_mt19937_64_standard_check_value():
    k = KATS["mt19937_64"]
    assert int(ob.mt19937_64(k["seed"], k["index"])[-1]) == k["value"]


def test_pixel_seeds_are_the_row_major_mt19937_64_stream():
    seeds = ob.seed_states(7, 5)
    assert np.array_equal(seeds, ob.mt19937_64(591842031321323413, 35))


# ---------------------------------------------------------------------------------------------
def test_cmf_table_matches_cie1931():
    cmf = cmf_table()
    assert cmf.shape == (3, 471)
    for lam, xyz in KATS["cie1931_2deg"].items():
        np.testing.assert_allclose(cmf[:, int(lam) - 360], xyz, rtol=0, atol=2e-6)
    sums = cmf.sum(axis=1)
    np.testing.assert_allclose(sums, [KATS["cie1931_sums"][c] for c in "xyz"], atol=2e-3)


def test_cmf_strata_integrate_the_table():
    cmf = cmf_table()
    lam = np.arange(360, 831, dtype=np.float64)
    bx, by, bz, integral = ob.cmf_bins()
    width = (830.0 - 360.0) / 16
    fine = np.linspace(360.0, 830.0, 16 * 4000 + 1)
    for c, bins in enumerate((bx, by, bz)):
        f = np.interp(fine, lam, cmf[c])
        for k in range(16):
            sl = slice(k * 4000, (k + 1) * 4000 + 1)
            exact = np.trapezoid(f[sl], fine[sl])
            assert abs(bins[k] - exact) <= 2e-4 * max(1.0, exact), (c, k, bins[k], exact)
    assert abs(integral - by.sum()) < 1e-4
    np.testing.assert_allclose(integral, KATS["cie1931_sums"]["y"], atol=2e-3)
    # equal-energy white has chromaticity (1/3, 1/3)
    X, Y, Z = bx.sum(), by.sum(), bz.sum()
    np.testing.assert_allclose([X / (X + Y + Z), Y / (X + Y + Z)], KATS["white_e"], atol=2e-4)
    assert width > 0


# ---------------------------------------------------------------------------------------------
REFLECTANCE, LIGHT_SOURCE, IOR = 0, 1, 2
REC709_LINEAR = 1


def upsampled_spectrum(kind, rgb):
    lam = np.arange(360, 831, dtype=np.float64)
    out = np.empty(len(lam))
    for i in range(0, len(lam), 4):
        l4 = np.full(4, lam[min(i + 3, len(lam) - 1)], np.float32)
        l4[:len(lam[i:i + 4])] = lam[i:i + 4]
        out[i:i + 4] = ob.meng_eval(kind, REC709_LINEAR, rgb, l4)[:len(lam[i:i + 4])]
    return out


@pytest.mark.parametrize("g", [0.05, 0.18, 0.5, 0.73, 1.0])
def test_meng_gray_reflectance_is_flat(g):
    s = upsampled_spectrum(REFLECTANCE, (g, g, g))
    np.testing.assert_allclose(s, g, rtol=1e-3)


@pytest.mark.parametrize("kind,white", [(REFLECTANCE, "white_e"), (LIGHT_SOURCE, "white_d65")])
def test_meng_spectrum_reproduces_tristimulus(kind, white):
    cmf = cmf_table()
    prim = [KATS["rec709_primaries"][c] for c in "rgb"]
    M = rgb_to_xyz_matrix(prim, KATS[white])
    rng = np.random.default_rng(11)
    for _ in range(40):
        rgb = rng.uniform(0.05, 1.0, 3)
        s = upsampled_spectrum(kind, tuple(rgb))
        xyz = (cmf * s).sum(axis=1) / cmf[1].sum()
        target = M @ rgb
        assert np.max(np.abs(xyz - target)) <= 2e-3 * target.max(), (rgb, xyz, target)


def test_meng_round_trip_through_rgb():
    rng = np.random.default_rng(5)
    for kind in (REFLECTANCE, LIGHT_SOURCE, IOR):
        for _ in range(20):
            rgb = rng.uniform(0.05, 1.0, 3)
            np.testing.assert_allclose(ob.triplet_rgb(kind, REC709_LINEAR, tuple(rgb)), rgb, atol=2e-3)
