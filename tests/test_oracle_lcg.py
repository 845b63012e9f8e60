"""The oracle's LCG (dsgl.hpp:453, :925, :1063, :1100) against the closed form
of the affine map s -> A*s + 11 mod 2^64, computed with Python integers."""
import random

from oracle import oracle as O

A = 25214903917
M = 1 << 64


def closed_form(s: int, n: int) -> int:
    # s_n = A^n s + 11 (A^n - 1)/(A - 1)  (mod 2^64); the geometric sum is
    # exact-divided in the integers before reduction.
    big = pow(A, n, M * (A - 1))
    geo = (big - 1) // (A - 1)
    return (pow(A, n, M) * s + 11 * geo) % M


def test_first_states_known_values():
    assert O.lcg_steps(0, 1) == 11
    assert O.lcg_steps(1, 1) == A + 11
    assert O.lcg_steps(1, 2) == ((A + 11) * A + 11) % M


def test_steps_match_closed_form():
    rnd = random.Random(5)
    for _ in range(50):
        s = rnd.getrandbits(64)
        n = rnd.choice([0, 1, 2, 3, 7, 64, 65, 1000, 4097])
        assert O.lcg_steps(s, n) == closed_form(s, n)


def test_jump_matches_sequential_and_closed_form():
    rnd = random.Random(7)
    for _ in range(50):
        s = rnd.getrandbits(64)
        n = rnd.randrange(0, 5000)
        assert O.lcg_jump(s, n) == O.lcg_steps(s, n)
    for n in (1 << 20, (1 << 40) + 12345, (1 << 63) + 1):
        s = rnd.getrandbits(64)
        assert O.lcg_jump(s, n) == closed_form(s, n)
