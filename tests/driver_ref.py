"""Pure-Python restatement of the search driver, used to pin the C++ driver
(graphgolf-cuda_amd/csrc/gg_search.cpp) move for move.  Test infrastructure.

RNG: std::mt19937_64 (the C++ standard's parameters: w=64, n=312, m=156,
r=31, a=0xB5026F5AA96619E9, u=29, d=0x5555555555555555, s=17,
b=0x71D67FFFEDA60000, t=37, c=0xFFF7EEE000000000, l=43, f=6364136223846793005),
Lemire's bounded integers, 53-bit doubles.  Acceptance: cppcode/annealing.cpp:
406-412 and cppcode/HillClimb.cpp:239.
"""
from __future__ import annotations

import math

import numpy as np

MASK = (1 << 64) - 1


class MT64:
    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & MASK
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & MASK
        self.idx = 312

    def _twist(self):
        mt = self.mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.idx = 0

    def next(self):
        if self.idx >= 312:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & MASK

    def below(self, rng):
        x = self.next()
        m = x * rng
        lo = m & MASK
        if lo < rng:
            t = ((1 << 64) - rng) % rng
            while lo < t:
                x = self.next()
                m = x * rng
                lo = m & MASK
        return m >> 64

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def propose(adj, rng):
    n, d = adj.shape
    a = rng.below(n); ia = rng.below(d)
    c = rng.below(n); ic = rng.below(d)
    kind = rng.below(2)
    b = int(adj[a, ia]); dd = int(adj[c, ic])
    if len({a, b, c, dd}) < 4:
        return None
    xa, xb = (dd, c) if kind else (c, dd)
    xc, xd = (b, a) if kind else (a, b)
    if xa in adj[a] or xb in adj[b]:
        return None
    ib = int(np.nonzero(adj[b] == a)[0][0]); idd = int(np.nonzero(adj[dd] == c)[0][0])
    return (a, b, c, dd, kind), [(a, ia, xa), (b, ib, xb), (c, ic, xc), (dd, idd, xd)]


def temperature(step, T0, T1, total_steps, schedule=0):
    """Temperature of step `step` (1-based) on the ASPL scale: geometric
    T0 -> T1, or the reference's T0 * 0.995^floor((step-1)/10^4)
    (cppcode/annealing.cpp:442, updated after every 10^4 iterations)."""
    if schedule == 1:
        return T0 * math.pow(0.995, (step - 1) // 10000)
    return T0 * math.pow(T1 / T0, step / total_steps)


def accept(rule, rng, step, NN1, x, y, T0, T1, total_steps, schedule):
    """Acceptance of y = (total, diameter, ok, wvc) against x = (total,
    diameter, wvc): cppcode/annealing.cpp:406-412, HillClimb.cpp:239,
    annealingWVC.cpp:412-417, HillClimbWVC.cpp:240.  Failed candidates carry
    the sentinels 100000000 / 1e8."""
    tx, dx, wx = x
    ty, dy, ok, wy = y
    fx = tx / NN1
    fy = ty / NN1 if ok else 1e8
    dyv = dy if ok else 100000000
    wyv = wy if ok else 100000000
    wvc = rule.endswith("_wvc")
    lex = ((dyv, wyv, fy) < (dx, wx, fx)) if wvc else ((dyv, fy) < (dx, fx))
    if rule.startswith("hill"):
        return ok and lex
    acc = False
    if (lex if wvc else (fy < fx or dyv < dx)):
        acc = True
    elif rng.uniform() < math.exp((fx - fy) * NN1 /
                                  (temperature(step, T0, T1, total_steps, schedule) * NN1)):
        acc = True
    if dyv > dx or (wvc and dyv == dx and wyv > wx) or not ok:
        acc = False
    return acc


def run(adj, evaluate, steps, mode_K=1, rule="anneal", T0=1e-3, T1=1e-6, total_steps=None,
        seed=12345, schedule=0):
    """evaluate(adj) -> (total_scaled, diameter, ok[, wvc]).  rule: anneal,
    hill, anneal_wvc, hill_wvc.  Returns the per-step log rows {step, a, b,
    c, d, kind, total, diameter, status, accepted} and the final graph."""
    adj = np.array(adj, dtype=np.int32)
    n, d = adj.shape
    total_steps = total_steps or steps
    rng = MT64(seed)
    ev = lambda g: (tuple(evaluate(g)) + (0,))[:4]  # noqa: E731
    tx, dx, _, wx = ev(adj)
    log = []
    NN1 = float(n) * (n - 1)
    wvc = rule.endswith("_wvc")
    for step in range(1, steps + 1):
        cands = []
        for _ in range(mode_K):
            while True:
                p = propose(adj, rng)
                if p is not None:
                    break
            cands.append(p)
        evals = []
        for mv, ents in cands:
            new = adj.copy()
            for v, s, x in ents:
                new[v, s] = x
            evals.append(ev(new) + (new,))
        # best by (ok, diameter, [wvc,] total)
        bi = min(range(len(evals)), key=lambda k: (not evals[k][2], evals[k][1],
                                                   evals[k][3] if wvc else 0, evals[k][0]))
        ty, dy, ok, wy, new = evals[bi]
        acc = accept(rule, rng, step, NN1, (tx, dx, wx), (ty, dy, ok, wy), T0, T1, total_steps,
                     schedule)
        mv = cands[bi][0]
        log.append([step, *mv, ty if ok else -1, dy if ok else -1, 0 if ok else 1, int(acc)])
        if acc:
            adj = new
            tx, dx, wx = ty, dy, wy
    return np.array(log, dtype=np.int64), adj


# ---- symmetric ("orbit") part graphs: cppcode/annealing.cpp:112-197 ----------------------

def _partner_slot(t, N, row, frm, slot, diff, n=None):
    for i in range(t.shape[1] if n is None else n):
        if row == frm and i == slot:
            continue
        if (int(t[row, i]) + diff) % N == 0:
            return i
    return -1


def _row_simple(t, N, r, n=None):
    xs = [int(x) % N for x in t[r][:n]]
    return 0 not in xs and len(set(xs)) == len(xs)


def propose_orbit(t, N, rng, lens=None):
    """One attempt; returns (move record, new table) or None (redraw).  With
    row lengths (annealingNR, cppcode/annealingNR.cpp:117,146-147) the slot is
    drawn among the row's entries and partners among the partner row's."""
    M, d = t.shape
    g = N // M
    ln = (lambda r: d) if lens is None else (lambda r: int(lens[r]))  # noqa: E731
    coin = rng.below(2)
    out = t.copy()
    if g > 1 and coin == 0:
        frm = rng.below(M)
        ifrm = rng.below(ln(frm))
        diff = int(t[frm, ifrm])
        to = ((frm + diff) % N) % M
        ito = _partner_slot(t, N, to, frm, ifrm, diff, ln(to))
        if ito < 0:
            return None
        if frm > to:
            frm, to, ifrm, ito = to, frm, ito, ifrm
        k = rng.below(g)
        nd = (to - frm) + M * k
        out[frm, ifrm] = nd
        out[to, ito] = -nd
        if not (_row_simple(out, N, frm, ln(frm)) and _row_simple(out, N, to, ln(to))):
            return None
        return (frm * d + ifrm, to * d + ito, k, 0, 0), out
    a = rng.below(M)
    ia = rng.below(ln(a))
    c = rng.below(M)
    ic = rng.below(ln(c))
    dab, dcd = int(t[a, ia]), int(t[c, ic])
    b = ((a + dab) % N) % M
    dd = ((c + dcd) % N) % M
    ib = _partner_slot(t, N, b, a, ia, dab, ln(b))
    idd = _partner_slot(t, N, dd, c, ic, dcd, ln(dd))
    if ib < 0 or idd < 0:
        return None
    if (c == a and ic == ia) or (c == b and ic == ib):
        return None
    ka, kb = rng.below(g), rng.below(g)
    rec = (a * d + ia, c * d + ic, ka, kb, 1)
    if a > dd:
        a, dd, ia, idd = dd, a, idd, ia
    out[a, ia] = (dd - a) + M * ka
    out[dd, idd] = -((dd - a) + M * ka)
    if b > c:
        b, c, ib, ic = c, b, ic, ib
    out[b, ib] = (c - b) + M * kb
    out[c, ic] = -((c - b) + M * kb)
    if not all(_row_simple(out, N, r, ln(r)) for r in (a, dd, b, c)):
        return None
    return rec, out


# ---- 2-D / 3-D symmetric graphs (piece, box): pieceannealing.cpp:113-204,
# boxHillClimb.cpp:113-220 -------------------------------------------------------------

class Lattice:
    """Per-axis orders Ns and base extents Ms; rows mixed-radix, axis 0 major."""

    def __init__(self, Ns, Ms, d):
        self.D, self.N, self.M, self.d = len(Ns), list(Ns), list(Ms), d
        self.g = [n // m for n, m in zip(Ns, Ms)]
        self.Mtot = int(np.prod(Ms))

    def coords(self, row):
        c = [0] * self.D
        for k in reversed(range(self.D)):
            c[k] = row % self.M[k]
            row //= self.M[k]
        return c

    def row_of(self, c):
        r = 0
        for k in range(self.D):
            r = r * self.M[k] + c[k]
        return r

    def target(self, row, off):
        c = self.coords(row)
        return self.row_of([((c[k] + int(off[k])) % self.N[k]) % self.M[k] for k in range(self.D)])

    def reverse(self, o, q):
        return all((int(o[k]) + int(q[k])) % self.N[k] == 0 for k in range(self.D))

    def partner(self, t, row, frm, slot, diff):
        for i in range(self.d):
            if row == frm and i == slot:
                continue
            if self.reverse(t[row, i], diff):
                return i
        return -1

    def row_simple(self, t, r):
        xs = [tuple(int(t[r, i, k]) % self.N[k] for k in range(self.D)) for i in range(self.d)]
        return all(any(x) for x in xs) and len(set(xs)) == len(xs)

    def row_symmetric(self, t, r):
        return all(self.partner(t, self.target(r, t[r, i]), r, i, t[r, i]) >= 0
                   for i in range(self.d))

    def lift(self, frm, to, rng):
        cf, ct = self.coords(frm), self.coords(to)
        out, kp = [], 0
        for k in range(self.D):
            kk = rng.below(self.g[k])
            out.append((ct[k] - cf[k]) + self.M[k] * kk)
            kp = kp * self.g[k] + kk
        return np.array(out, dtype=np.int32), kp


def propose_lattice(t, L, rng):
    """One attempt on a [Mtot, d, D] table; (move record, new table) or None."""
    d = L.d
    lengths = all(g > 1 for g in L.g)
    coin = rng.below(2)
    out = t.copy()
    if lengths and coin == 0:
        frm, ifrm = rng.below(L.Mtot), rng.below(d)
        diff = t[frm, ifrm].copy()
        to = L.target(frm, diff)
        ito = L.partner(t, to, frm, ifrm, diff)
        if ito < 0:
            return None
        nd, kp = L.lift(frm, to, rng)
        out[frm, ifrm] = nd
        out[to, ito] = -nd
        if not all(L.row_simple(out, r) and L.row_symmetric(out, r) for r in (frm, to)):
            return None
        return (frm * d + ifrm, to * d + ito, kp, 0, 0), out
    a, ia = rng.below(L.Mtot), rng.below(d)
    c, ic = rng.below(L.Mtot), rng.below(d)
    dab, dcd = t[a, ia].copy(), t[c, ic].copy()
    b, dd = L.target(a, dab), L.target(c, dcd)
    ib = L.partner(t, b, a, ia, dab)
    idd = L.partner(t, dd, c, ic, dcd)
    if ib < 0 or idd < 0:
        return None
    if (c == a and ic == ia) or (c == b and ic == ib):
        return None
    dad, ka = L.lift(a, dd, rng)
    dbc, kb = L.lift(b, c, rng)
    out[a, ia] = dad
    out[dd, idd] = -dad
    out[b, ib] = dbc
    out[c, ic] = -dbc
    if not all(L.row_simple(out, r) and L.row_symmetric(out, r) for r in (a, dd, b, c)):
        return None
    return (a * d + ia, c * d + ic, ka, kb, 1), out


def random_lattice_table(Ns, Ms, d, seed):
    """Seeded simple symmetric table [Mtot, d, D]: the Mtot*d stubs paired at
    random, each pair an edge orbit with a uniform lift per axis, resampled
    until no self-loop, half-orbit or repeated offset remains."""
    L = Lattice(Ns, Ms, d)
    rng = np.random.default_rng(seed)
    for _ in range(10000):
        stubs = rng.permutation(L.Mtot * d)
        t = np.zeros((L.Mtot, d, L.D), dtype=np.int32)
        fill = [0] * L.Mtot
        for i in range(0, len(stubs), 2):
            r1, r2 = int(stubs[i]) // d, int(stubs[i + 1]) // d
            c1, c2 = L.coords(r1), L.coords(r2)
            off = np.array([(c2[k] - c1[k]) + L.M[k] * int(rng.integers(0, L.g[k]))
                            for k in range(L.D)], dtype=np.int32)
            t[r1, fill[r1]] = off
            fill[r1] += 1
            t[r2, fill[r2]] = -off
            fill[r2] += 1
        if all(L.row_simple(t, r) and L.row_symmetric(t, r) for r in range(L.Mtot)):
            return t
    raise RuntimeError("no simple table found")


def irregular_part(t, N, drop, seed):
    """A symmetric part table whose rows have different lengths (annealingNR's
    input): `drop` edge orbits of t removed (an entry and its reverse in the
    partner row), each row's remaining entries moved to its front, padding
    offset 0.  Returns (table, row lengths)."""
    rng = np.random.default_rng(seed)
    M, d = t.shape
    rows = [[int(x) for x in r] for r in t]
    removed = 0
    while removed < drop:
        r = int(rng.integers(M))
        if len(rows[r]) <= 1:
            continue
        i = int(rng.integers(len(rows[r])))
        x = rows[r][i]
        p = ((r + x) % N) % M
        js = [k for k in range(len(rows[p])) if (p != r or k != i) and (rows[p][k] + x) % N == 0]
        if not js or len(rows[p]) <= 1:
            continue
        if p == r:
            for k in sorted((i, js[0]), reverse=True):
                del rows[r][k]
        else:
            del rows[r][i]
            del rows[p][js[0]]
        removed += 1
    lens = np.array([len(r) for r in rows], dtype=np.int32)
    out = np.zeros((M, d), dtype=np.int32)
    for r, row in enumerate(rows):
        out[r, :len(row)] = row
    return out, lens


def run_orbit(N, table, evaluate, steps, rule="hill", T0=1e-3, T1=1e-6, total_steps=None,
              seed=1, propose=None, lens=None):
    """evaluate(table) -> (total_scaled, diameter, ok[, wvc]).  Returns the
    per-step log rows {step, a*d+slot_a, c*d+slot_c, k1, k2, kind, total,
    diameter, status, accepted} and the final table."""
    t = np.array(table, dtype=np.int32)
    total_steps = total_steps or steps
    rng = MT64(seed)
    ev = lambda g: (tuple(evaluate(g)) + (0,))[:4]  # noqa: E731
    tx, dx, _, wx = ev(t)
    NN1 = float(N) * (N - 1)
    log = []
    for step in range(1, steps + 1):
        while True:
            p = propose(t, rng) if propose else propose_orbit(t, N, rng, lens)
            if p is not None:
                break
        rec, new = p
        ty, dy, ok, wy = ev(new)
        acc = accept(rule, rng, step, NN1, (tx, dx, wx), (ty, dy, ok, wy), T0, T1, total_steps, 0)
        log.append([step, *rec, ty if ok else -1, dy if ok else -1, 0 if ok else 1, int(acc)])
        if acc:
            t = new
            tx, dx, wx = ty, dy, wy
    return np.array(log, dtype=np.int64), t
