"""Small seeded sequence generators for tests (numpy Generator, fixed seeds)."""
import numpy as np

ACGT = np.frombuffer(b"ACGT", np.uint8)


def rand_codes(rng, n, p_n=0.0, alphabet=4):
    s = rng.integers(0, alphabet, n).astype(np.uint8)
    if p_n:
        s[rng.random(n) < p_n] = 4
    return s


def mutate_codes(rng, s, p_sub, p_indel, max_indel=5):
    out = []
    i = 0
    n = len(s)
    while i < n:
        u = rng.random()
        if u < p_sub:
            b = int(s[i])
            out.append((b + 1 + int(rng.integers(0, 3))) % 4 if b < 4 else b)
            i += 1
        elif u < p_sub + p_indel / 2:
            out.extend(rng.integers(0, 4, int(rng.integers(1, max_indel + 1))).tolist())
        elif u < p_sub + p_indel:
            i += int(rng.integers(1, max_indel + 1))
        else:
            out.append(int(s[i]))
            i += 1
    return np.array(out, np.uint8)


def codes_to_ascii(c):
    lut = np.frombuffer(b"ACGTN", np.uint8)
    return lut[np.asarray(c, np.uint8)].tobytes()
