"""Writes rng_streams.json: the random streams of the reference's per-pixel
RNG (include/core/rng.cuh:8-51, src/core/rng.cu:5-41) for fixed
(pixel, sample, strategy) keys.

The generators are implemented here from their published definitions and
first checked against the published known answers of those definitions:
  * PCG32 (XSH-RR 64/32, O'Neill): pcg32_srandom(42, 54) produces
    0xa15c02b7 0x7b47f409 0xba1d3330 0x83d2f293 0xbfa4784b 0xcbed606e
    (pcg-c-basic demo output);
  * Marsaglia xorshift32 (13, 17, 5): from 2463534242 the first output is
    723471715 (Marsaglia, "Xorshift RNGs", 2003);
  * SplitMix64 (Steele, Lea, Flood / Vigna's splitmix64.c): from state
    1234567 the outputs are 6457827717110365317, 3203168211198807973,
    9817491932198370423, 4593380528125082431, 16408922859458223821.
The reference's use of them (keys, SplitMix seeding, the fixed PCG
increment, the 24-bit float mapping) is then applied.  Run:
    python tests/golden/make_rng_golden.py
"""
import json
import os

import numpy as np

M64 = (1 << 64) - 1
M32 = (1 << 32) - 1


def splitmix_next(state):
    state = (state + 0x9E3779B97F4A7C15) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


def pcg_step(state, inc):
    old = state
    state = (old * 6364136223846793005 + inc) & M64
    xorshifted = (((old >> 18) ^ old) >> 27) & M32
    rot = old >> 59
    return state, ((xorshifted >> rot) | (xorshifted << ((32 - rot) & 31))) & M32


def xorshift32(x):
    x ^= (x << 13) & M32
    x ^= x >> 17
    x ^= (x << 5) & M32
    return x & M32


def self_check():
    inc = ((54 << 1) | 1) & M64
    state, _ = pcg_step(0, inc)
    state = (state + 42) & M64
    state, _ = pcg_step(state, inc)
    out = []
    for _ in range(6):
        state, v = pcg_step(state, inc)
        out.append(v)
    assert out == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E], [hex(v) for v in out]
    assert xorshift32(2463534242) == 723471715
    s, outs = 1234567, []
    for _ in range(5):
        s, v = splitmix_next(s)
        outs.append(v)
    assert outs == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                    16408922859458223821], outs


def stream(pixel, sample, strategy, count):
    """Reference key -> seed -> stream of 24-bit floats."""
    if strategy == 0:
        key = (pixel ^ ((sample * 0x9E3779B9) & M32)) & M32
        seed = (key * 747796405 + 2891336453) & M32
        state = 1 if seed == 0 else seed
    else:
        key = ((sample << 32) | pixel) & M64
        _, state = splitmix_next(key)
    seeded = state
    values = []
    for _ in range(count):
        if strategy == 0:
            state = xorshift32(state)
            v = state
        else:
            state, v = pcg_step(state, 1442695040888963407)
        values.append(float(np.float32(v >> 8) * np.float32(2.0 ** -24)))
    return seeded, values


def main():
    self_check()
    cases = []
    for strategy in (0, 1):
        for pixel, sample in [(0, 0), (1, 0), (0, 1), (1234, 567), (2073599, 63), (0xFFFFFFFF, 0xFFFFFFFF)]:
            seeded, values = stream(pixel, sample, strategy, 16)
            cases.append({"pixel": pixel, "sample": sample, "strategy": strategy, "seeded_state": seeded,
                          "floats": values})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_streams.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/make_rng_golden.py", "cases": cases}, f, indent=1)
    print(f"wrote {path} ({len(cases)} cases)")


if __name__ == "__main__":
    main()
