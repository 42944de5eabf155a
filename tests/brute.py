"""Brute-force enumeration of alignments on tiny inputs (test helper).

Enumerates every alignment path explicitly and scores it from the gap
model's definition -- no dynamic programming -- so it checks the oracle's
recurrences, borders and seeds independently.  Inputs used with it are
integer-valued, so every sum is exact in any order.
"""

import itertools


def global_paths(n, m):
    """All move strings from (0,0) to (n,m): 'D' diag, 'U' consumes s, 'L' consumes t."""
    out = []

    def rec(i, j, acc):
        if i == n and j == m:
            out.append("".join(acc))
            return
        if i < n and j < m:
            acc.append("D"); rec(i + 1, j + 1, acc); acc.pop()
        if i < n:
            acc.append("U"); rec(i + 1, j, acc); acc.pop()
        if j < m:
            acc.append("L"); rec(i, j + 1, acc); acc.pop()

    rec(0, 0, [])
    return out


def run_cost(gap, k):
    """Cost of a maximal gap run of length k (pyalign/gaps.py definitions)."""
    if gap[0] == "linear":
        return gap[1] * k
    open_, extend = gap[1], gap[2]
    return open_ + extend * (k - 1)  # w(k) = u + v k with u = open - extend, v = extend


def score_path(moves, a, b, S, gap, maximize):
    """Score of a move string starting at s[a], t[b]; returns (score, edges)."""
    i, j = a, b
    total = 0.0
    edges = []
    sgn = -1.0 if maximize else 1.0
    for mv, grp in itertools.groupby(moves):
        k = len(list(grp))
        if mv == "D":
            for _ in range(k):
                total += S[i][j]
                edges.append((i, j))
                i += 1
                j += 1
        elif mv == "U":
            total += sgn * run_cost(gap, k)
            i += k
        else:
            total += sgn * run_cost(gap, k)
            j += k
    return total, tuple(edges)


def best(values, maximize):
    return max(values) if maximize else min(values)


def brute_global(S, gap, maximize=True):
    n, m = len(S), len(S[0]) if len(S) else 0
    scored = [score_path(p, 0, 0, S, gap, maximize) for p in global_paths(n, m)]
    opt = best([x for x, _ in scored], maximize)
    return opt, sorted(set(e for x, e in scored if x == opt))


def brute_local(S, gap, maximize=True):
    """Smith-Waterman optimum: best over all substring pairs, floored at 0."""
    n, m = len(S), len(S[0])
    vals = [0.0]
    for a in range(n):
        for b in range(m):
            for c in range(a + 1, n + 1):
                for d in range(b + 1, m + 1):
                    sub = [row[b:d] for row in S[a:c]]
                    for p in global_paths(c - a, d - b):
                        vals.append(score_path(p, 0, 0, sub, gap, maximize)[0])
    return best(vals, maximize)


def brute_semiglobal(S, gap, maximize=True):
    """Free leading gaps on one side (start on the top or left border) and
    free trailing gaps (end on the last row or the last column)."""
    n, m = len(S), len(S[0])
    vals = []
    for a in range(n + 1):
        for b in range(m + 1):
            if not (a == 0 or b == 0):
                continue
            for c in range(a, n + 1):
                for d in range(b, m + 1):
                    if not (c == n or d == m):
                        continue
                    if c < 1 or d < 1 or (c == a and d == b):
                        continue
                    sub = [row[b:d] for row in S[a:c]]
                    for p in global_paths(c - a, d - b):
                        vals.append(score_path(p, 0, 0, sub, gap, maximize)[0])
    return best(vals, maximize)


def dtw_paths(n, m):
    out = []

    def rec(i, j, acc):
        acc.append((i, j))
        if i == n - 1 and j == m - 1:
            out.append(list(acc))
        else:
            if i + 1 < n and j + 1 < m:
                rec(i + 1, j + 1, acc)
            if i + 1 < n:
                rec(i + 1, j, acc)
            if j + 1 < m:
                rec(i, j + 1, acc)
        acc.pop()

    rec(0, 0, [])
    return out


def brute_dtw(C, maximize=False):
    n, m = len(C), len(C[0])
    vals = [sum(C[i][j] for i, j in p) for p in dtw_paths(n, m)]
    return best(vals, maximize)


# ---- general gap costs (Waterman-Smith-Beyer) --------------------------
# A gap is a jump of any length k >= 1 costing w(k); consecutive jumps in
# the same direction are separate gaps (the recurrence allows them), so
# paths are enumerated as jump sequences, not as maximal runs.  The border
# row and column are fixed values, not computed by the recurrence: a path
# that leaves the origin along a border does so in one jump.

def jump_paths(n, m):
    """All jump sequences from (0,0) to (n,m): ('D', 1), ('U', k), ('L', k);
    along row 0 or column 0 only one jump, from the origin."""
    out = []

    def rec(i, j, acc):
        if i == n and j == m:
            out.append(tuple(acc))
            return
        if i < n and j < m:
            acc.append(("D", 1)); rec(i + 1, j + 1, acc); acc.pop()
        if j > 0 or i == 0:
            for k in range(1, n - i + 1):
                acc.append(("U", k)); rec(i + k, j, acc); acc.pop()
        if i > 0 or j == 0:
            for k in range(1, m - j + 1):
                acc.append(("L", k)); rec(i, j + k, acc); acc.pop()

    rec(0, 0, [])
    return out


def score_jumps(jumps, a, b, S, ws, wt, maximize):
    i, j = a, b
    total = 0.0
    edges = []
    sgn = -1.0 if maximize else 1.0
    for mv, k in jumps:
        if mv == "D":
            total += S[i][j]
            edges.append((i, j))
            i += 1
            j += 1
        elif mv == "U":
            total += sgn * ws[k]
            i += k
        else:
            total += sgn * wt[k]
            j += k
    return total, tuple(edges)


def brute_wsb_global(S, ws, wt, maximize=True):
    n, m = len(S), len(S[0])
    scored = [score_jumps(p, 0, 0, S, ws, wt, maximize) for p in jump_paths(n, m)]
    opt = best([x for x, _ in scored], maximize)
    return opt, sorted(set(e for x, e in scored if x == opt))


def brute_wsb_local(S, ws, wt):
    """Maximisation: the best jump path over all substring pairs, floored at 0."""
    n, m = len(S), len(S[0])
    vals = [0.0]
    for a in range(n):
        for b in range(m):
            for c in range(a + 1, n + 1):
                for d in range(b + 1, m + 1):
                    sub = [row[b:d] for row in S[a:c]]
                    for p in jump_paths(c - a, d - b):
                        vals.append(score_jumps(p, 0, 0, sub, ws, wt, True)[0])
    return max(vals)


def brute_wsb_semiglobal(S, ws, wt):
    """Maximisation: start on the top or left border, end on the last row or
    column, jump paths in between."""
    n, m = len(S), len(S[0])
    vals = []
    for a in range(n + 1):
        for b in range(m + 1):
            if not (a == 0 or b == 0):
                continue
            for c in range(a, n + 1):
                for d in range(b, m + 1):
                    if not (c == n or d == m):
                        continue
                    if c < 1 or d < 1 or (c == a and d == b):
                        continue
                    sub = [row[b:d] for row in S[a:c]]
                    for p in jump_paths(c - a, d - b):
                        vals.append(score_jumps(p, 0, 0, sub, ws, wt, True)[0])
    return max(vals)
