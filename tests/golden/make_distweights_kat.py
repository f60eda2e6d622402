"""Writes distweights_kat.json: the known-answer cases of the reference's
test/source.jl:17-133 (inputs and the expected (indices, weights) written in
closed form there), evaluated with plain Python floats.

Run from the repository root: python tests/golden/make_distweights_kat.py
"""
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    dlg = list(range(2, 23, 2))                       # source.jl:18  ∆lg = 2:2:22
    lp_ = [-1]
    for d in dlg:
        lp_.append(lp_[-1] + d)                       # :19  l′ = cumsum([-1; ∆lg])
    lprim_g = [(a + b) / 2 for a, b in zip(lp_[:-1], lp_[1:])]   # :21 movingavg
    ldual_g = lp_[:-1]                                # :22
    domain = [lprim_g[0], lprim_g[-1]]                # :24
    lprim, dlprim = lprim_g[:-1], [b - a for a, b in zip(ldual_g[:-1], ldual_g[1:])]   # :26
    ldual, dldual = ldual_g[1:], [b - a for a, b in zip(lprim_g[:-1], lprim_g[1:])]    # :27
    P, D = 0, 1
    dp = lambda i: 1 / dlprim[i - 1]
    dd = lambda i: 1 / dldual[i - 1]
    cases = []

    def add(line, c, gt, bloch, ind=None, wt=None, throws=False):
        cases.append(dict(line=line, c=c, gt=gt, isbloch=bloch, throws=throws, ind=ind, wt=wt))

    for b in (True, False):                                        # :30-32
        add(31 if b else 32, 8, P, b, [3, 3], [dp(3), 0])
    r = (11 - 8) / (15 - 8)                                        # :34-37
    for b in (True, False):
        add(36 if b else 37, 11, P, b, [3, 4], [dp(3) * (1 - r), dp(4) * r])
    for b in (True, False):                                        # :39-41
        add(40 if b else 41, 3, P, b, [2, 2], [dp(2), 0])
    r = (3 - 1) / (3 - 0)                                          # :43-46
    add(45, 1, P, True, [1, 2], [dp(1) * r, dp(2) * (1 - r)])
    add(46, 1, P, False, [2, 2], [dp(2) * (1 - r), 0])
    add(49, 0, P, True, [1, 1], [dp(1), 0])                        # :48-50
    add(50, 0, P, False, [2, 2], [0, 0])
    for c, ln in ((-0.5, 53), (-1, 57)):                           # :52-58
        add(ln, c, P, True, throws=True)
        add(ln + 1, c, P, False, throws=True)
    for b in (True, False):                                        # :60-62
        add(61 if b else 62, 99, P, b, [10, 10], [dp(10), 0])
    r = (105 - 99) / (120 - 99)                                    # :64-67
    add(66, 105, P, True, [10, 1], [dp(10) * (1 - r), dp(1) * r])
    add(67, 105, P, False, [10, 10], [dp(10) * (1 - r), 0])
    r = (120 - 99) / (120 - 99)                                    # :69-72
    add(71, 120, P, True, [10, 1], [dp(10) * (1 - r), dp(1) * r])
    add(72, 120, P, False, [10, 10], [dp(10) * (1 - r), 0])

    for b in (True, False):                                        # :77-79
        add(78 if b else 79, 11, D, b, [3, 3], [dd(3), 0])
    r = (15 - 11) / (19 - 11)                                      # :81-84
    for b in (True, False):
        add(83 if b else 84, 15, D, b, [3, 4], [dd(3) * (1 - r), dd(4) * r])
    for b in (True, False):                                        # :86-88
        add(87 if b else 88, 1, D, b, [1, 1], [dd(1), 0])
    r = (1 - 0.5) / ((1 - 0) + (120 - 109))                        # :90-93
    add(92, 0.5, D, True, [1, 10], [dd(1) * (1 - r), dd(10) * r])
    add(93, 0.5, D, False, [1, 1], [dd(1), 0])
    r = (1 - 0) / ((1 - 0) + (120 - 109))                          # :95-98
    add(97, 0, D, True, [1, 10], [dd(1) * (1 - r), dd(10) * r])
    add(98, 0, D, False, [1, 1], [dd(1), 0])
    add(101, -0.5, D, True, throws=True)                           # :100-102
    add(102, -0.5, D, False, throws=True)
    for b in (True, False):                                        # :104-106
        add(105 if b else 106, 109, D, b, [10, 10], [dd(10), 0])
    r = (115 - 109) / ((120 - 109) + (1 - 0))                      # :108-111
    add(110, 115, D, True, [10, 1], [dd(10) * (1 - r), dd(1) * r])
    add(111, 115, D, False, [10, 10], [dd(10), 0])
    r = (120 - 109) / ((120 - 109) + (1 - 0))                      # :113-116
    add(115, 120, D, True, [10, 1], [dd(10) * (1 - r), dd(1) * r])
    add(116, 120, D, False, [10, 10], [dd(10), 0])

    grid = dict(domain=domain, lprim=lprim, dlprim=dlprim, ldual=ldual, dldual=dldual)
    # N = 1 cases (source.jl:118-132): seeded instead of the reference's rand()
    rng = random.Random(20240601)
    n1 = dict(domain=[0, 2], lprim=[0], ldual=[1], dlprim=[9], dldual=[10],
              c=[2 * rng.random() for _ in range(4)], sum_prim=1 / 9, sum_dual=1 / 10)
    with open(os.path.join(HERE, "distweights_kat.json"), "w") as f:
        json.dump(dict(source="test/source.jl:17-133", grid=grid, cases=cases, n1=n1), f, indent=1)


if __name__ == "__main__":
    main()
