"""Writes published_kats.json: known-answer values from published standards and reference
implementations, used to pin the oracle.  Nothing here is computed by this project.

Sources
  pcg32_demo      pcg-c-basic, pcg32-demo (M. O'Neill, 2014): pcg32_srandom_r(42, 54), then six
                  pcg32_random_r outputs.  srandom sets inc = (54 << 1) | 1, state = 0, steps,
                  adds 42, steps.
  mt19937_64      ISO C++ [rand.predef]: the 10000th consecutive output of a default-constructed
                  std::mt19937_64 (seed 5489) is 9981545732273789042.
  cie1931_2deg    CIE 015:2018 / ISO 11664-1 colour-matching functions of the 1931 2-degree
                  standard observer, selected wavelengths (nm): xbar, ybar, zbar.
  cie1931_sums    Sums of the 1 nm tables over 360-830 nm.
  rec709_primaries, white points  ITU-R BT.709 primaries; D65 and equal-energy (E) white
                  chromaticities.
  fresnel_n15_normal  ((n - 1) / (n + 1))^2 for n = 1.5.
"""
import json
import os

KATS = {
    "pcg32_demo": {"initstate": 42, "initseq": 54,
                   "outputs": [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]},
    "mt19937_64": {"seed": 5489, "index": 10000, "value": 9981545732273789042},
    "cie1931_2deg": {
        "380": [0.001368, 0.000039, 0.006450], "400": [0.014310, 0.000396, 0.067850],
        "420": [0.134380, 0.004000, 0.645600], "440": [0.348280, 0.023000, 1.747060],
        "460": [0.290800, 0.060000, 1.669200], "480": [0.095640, 0.139020, 0.812950],
        "500": [0.004900, 0.323000, 0.272000], "520": [0.063270, 0.710000, 0.078250],
        "540": [0.290400, 0.954000, 0.020300], "555": [0.512050, 1.000000, 0.005750],
        "560": [0.594500, 0.995000, 0.003900], "580": [0.916300, 0.870000, 0.001650],
        "600": [1.062200, 0.631000, 0.000800], "620": [0.854450, 0.381000, 0.000190],
        "640": [0.447900, 0.175000, 0.000020], "660": [0.164900, 0.061000, 0.000000],
        "680": [0.046770, 0.017000, 0.000000], "700": [0.011359, 0.004102, 0.000000]},
    "cie1931_sums": {"x": 106.865, "y": 106.857, "z": 106.892},
    "rec709_primaries": {"r": [0.64, 0.33], "g": [0.30, 0.60], "b": [0.15, 0.06]},
    "white_d65": [0.3127, 0.3290],
    "white_e": [1.0 / 3.0, 1.0 / 3.0],
    "fresnel_n15_normal": 0.04,
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "published_kats.json")
    with open(out, "w") as f:
        json.dump(KATS, f, indent=1)
        f.write("\n")
    print("wrote", out)
