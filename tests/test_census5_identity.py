"""The closed forms behind the P2, P3, P5 and P6 census (census.cc,
census5_kernels.hip), evaluated term by term in plain Python on small
random graphs: every accumulator the device computes is restated here
twice -- from first principles (sets) and in the device's decomposition
over the degree-oriented DAG (anchor = the lowest-ranked vertex of each
occurrence) -- and the patterns they give are checked against
brute-force induced-subgraph counts (oracle.brute) and the oracle's
literal restatement of the reference's set algebra:

  P2 = induced forks, P3 = 4 x induced K1,4 + 3 x induced claws,
  P5 = 2 x induced bulls, P6 = 2 x induced (diamond + pendant at a
  degree-3 vertex) + 2 x induced paws.

Induced counts follow from subgraph (not necessarily induced) counts N(H)
by inverting the containment matrix of the 21 connected graphs on five
vertices; the rows used are written out in `_patterns`, and the test also
rebuilds that matrix by brute force and checks those rows against its
inverse."""
import itertools
import random
from math import comb

import numpy as np
import pytest

import oracle
from conftest import csr_from_edges


def _graph(seed):
    rng = random.Random(seed)
    n = rng.randint(6, 18)
    p = rng.choice([0.2, 0.35, 0.5, 0.7, 0.9])
    edges = [(a, b) for a in range(n) for b in range(a + 1, n)
             if rng.random() < p]
    adj = [set() for _ in range(n)]
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    return n, edges, adj


def _first_principles(adj):
    """Subgraph counts N(H) straight from their definitions."""
    n = len(adj)
    d = [len(s) for s in adj]
    E = [(a, b) for a in range(n) for b in adj[a] if a < b]
    t = {}
    for a, b in E:
        t[(a, b)] = t[(b, a)] = len(adj[a] & adj[b])
    T = [(a, b, c) for a, b in E for c in adj[a] & adj[b] if c > b]
    kT = {tr: len(adj[tr[0]] & adj[tr[1]] & adj[tr[2]]) for tr in T}
    K4 = sum(kT.values()) // 4
    K5 = 0
    for a, b, c in T:
        com = sorted(x for x in adj[a] & adj[b] & adj[c] if x > c)
        K5 += sum(1 for i, x in enumerate(com) for y in com[i + 1:]
                  if y in adj[x])
    rot = lambda tr: ((tr[0], tr[1], tr[2]), (tr[1], tr[0], tr[2]),
                      (tr[2], tr[0], tr[1]))
    tv = [sum(t[(v, u)] for u in adj[v]) // 2 for v in range(n)]
    s = [sum(d[x] - 1 for x in adj[v]) for v in range(n)]
    cod = {(a, c): adj[a] & adj[c] for a in range(n) for c in range(a + 1, n)}

    def e_in(S):
        return sum(1 for x in S for y in adj[x] & S if x < y)

    cyc = []  # 4-cycles a-b-c-e-a, each once
    for a, b, c, e in itertools.permutations(range(n), 4):
        if (a < b and a < c and a < e and b < e and b in adj[a]
                and c in adj[b] and e in adj[c] and a in adj[e]):
            cyc.append((a, b, c, e))
    q = {}
    for cy in cyc:
        for i in range(4):
            x, y = cy[i], cy[(i + 1) % 4]
            q[(x, y)] = q.get((x, y), 0) + 1
            q[(y, x)] = q.get((y, x), 0) + 1
    pairs_t = sum(comb(t[e], 2) for e in E)

    def c4_in(H):
        return sum(comb(len(adj[x] & adj[y] & H), 2)
                   for x, y in itertools.combinations(sorted(H), 2)) // 2

    N = {
        "star": sum(comb(x, 4) for x in d),
        "fork": sum(comb(d[c] - 1, 2) * (d[x] - 1) - t[(c, x)] * (d[c] - 2)
                    for a, b in E for c, x in ((a, b), (b, a))),
        "cricket": sum(comb(d[v] - 2, 2) for tr in T for v in tr),
        "bull": sum((d[a] - 2) * (d[b] - 2) - (t[(a, b)] - 1)
                    for tr in T for a, b in itertools.combinations(tr, 2)),
        "lolli": sum(s[a] - d[b] - d[c] + 2 - (t[(a, b)] - 1)
                     - (t[(a, c)] - 1) for tr in T for a, b, c in rot(tr)),
        "banner": sum(d[v] - 2 - (1 if cy[(i + 2) % 4] in adj[v] else 0)
                      for cy in cyc for i, v in enumerate(cy)),
        "dp3": sum(comb(t[e], 2) * (d[e[0]] + d[e[1]] - 6) for e in E),
        "bowtie": sum(comb(tv[v], 2) - sum(comb(t[(v, u)], 2)
                                           for u in adj[v])
                      for v in range(n)),
        "dp2": sum((t[e] - 1) * (d[c] - 2) for e in E
                   for c in adj[e[0]] & adj[e[1]]) - 12 * K4,
        "house": sum(t[e] * q.get(e, 0) for e in E) - 4 * pairs_t,
        "k23": sum(comb(len(S), 3) for S in cod.values()),
        "book3": sum(comb(t[e], 3) for e in E),
        "k4pend": sum(kT[tr] * (d[tr[0]] + d[tr[1]] + d[tr[2]] - 9)
                      for tr in T) // 3,
        "gem": sum((t[(h, x)] - 1) * (t[(h, y)] - 1)
                   for tr in T for h, x, y in rot(tr)) - 12 * K4,
        "k23e": sum(e_in(S) * (len(S) - 2) for S in cod.values()),
        "k4wing": sum(kT[tr] * (t[(tr[0], tr[1])] + t[(tr[0], tr[2])]
                                + t[(tr[1], tr[2])] - 6) for tr in T) // 2,
        "w4": sum(c4_in(adj[h]) for h in range(n)),
        "k5e": sum(comb(k, 2) for k in kT.values()),
        "k5": K5,
        "paw": sum(d[a] + d[b] + d[c] - 6 for a, b, c in T),
        "pairs_t": pairs_t, "k4": K4,
        "claw_sums": (sum(comb(x, 3) for x in d)
                      - sum(d[a] + d[b] + d[c] - 6 for a, b, c in T)),
    }
    return N


def _dag_decomposition(adj):
    """W4, K2,3, K2,3 + edge and the 4-cycle terms as the device computes
    them: anchors are the lowest-ranked vertex of each occurrence in the
    degree order of graph.cc:205-235 (degree, then id)."""
    n = len(adj)
    d = [len(s) for s in adj]
    up = lambda a, b: (d[b], b) > (d[a], a)
    Np = [{b for b in adj[a] if up(a, b)} for a in range(n)]
    Nm = [{b for b in adj[a] if up(b, a)} for a in range(n)]
    t = {(a, b): len(adj[a] & adj[b]) for a in range(n) for b in adj[a]}
    w4_out = w4_rim = k23 = a16 = b16 = 0
    table = {}
    # anchor pass: masks M_a[c] = N+(a) ∩ N(c) and lo_a[c] for every c
    # above a; codeg(a, c) = lo + |M|; pairs with lo >= 1 go to the table
    for a in range(n):
        cand = set()
        for y in Np[a]:
            cand |= {c for c in adj[y] if up(a, c)}
        lo = {}
        for z in Nm[a]:
            for c in Np[z]:
                if up(a, c):
                    lo[c] = lo.get(c, 0) + 1
                    cand.add(c)
        for c in cand:
            M = Np[a] & adj[c]
            cd = lo.get(c, 0) + len(M)
            deg = [len(adj[p] & M) for p in M]
            k23 += comb(cd, 3)
            w4_rim += sum(comb(x, 2) for x in deg)
            a16 += (sum(deg) // 2) * (cd - 2)
            if lo.get(c, 0):
                table[(a, c)] = cd
    # out-neighbourhood pass: 4-cycles inside G[N+(x)], and diamonds whose
    # lowest vertex is a hinge endpoint x (tips p, q in N+(x))
    for x in range(n):
        P = sorted(Np[x])
        acc = 0
        for p, q in itertools.combinations(P, 2):
            h = len(Np[x] & adj[p] & adj[q])
            acc += comb(h, 2)
            if h:
                lo_v, hi_v = (p, q) if up(p, q) else (q, p)
                b16 += h * (table[(lo_v, hi_v)] - 2)
        assert acc % 2 == 0
        w4_out += acc // 2
    # oriented 4-cycles: top v, opposite w, lower middles u
    c4 = tq = ban = 0
    for v in range(n):
        cnt, sdeg = {}, {}
        for u in Nm[v]:
            for w in adj[u]:
                if w != v and up(w, v):
                    cnt[w] = cnt.get(w, 0) + 1
                    sdeg[w] = sdeg.get(w, 0) + d[u] - 2
        for w, c in cnt.items():
            c4 += comb(c, 2)
            ban += comb(c, 2) * (d[v] + d[w] - 4) + (c - 1) * sdeg[w]
        for u in Nm[v]:
            for w in adj[u]:
                if w != v and up(w, v):
                    tq += (cnt[w] - 1) * (t[(v, u)] + t[(u, w)])
    return {"w4": w4_out + w4_rim, "k23": k23, "k23e": a16 + b16, "c4": c4,
            "tq": tq, "ban": ban}


def _patterns(N):
    """Induced counts from the subgraph counts (rows of the inverted
    containment matrix), then the reference's multiplicities."""
    g = N.__getitem__
    star = (g("star") - g("cricket") + g("dp3") + g("bowtie") - 2 * g("book3")
            - g("k4pend") - g("gem") + 2 * g("k4wing") + g("w4")
            - 3 * g("k5e") + 5 * g("k5"))
    fork = (g("fork") - 2 * g("cricket") - 2 * g("bull") - g("lolli")
            - 2 * g("banner") + 5 * g("dp3") + 4 * g("bowtie") + 4 * g("dp2")
            + 4 * g("house") + 6 * g("k23") - 12 * g("book3")
            - 9 * g("k4pend") - 10 * g("gem") - 10 * g("k23e")
            + 20 * g("k4wing") + 20 * g("w4") - 36 * g("k5e") + 60 * g("k5"))
    bull = (g("bull") - 2 * g("dp3") - 2 * g("dp2") - g("house")
            + 6 * g("book3") + 6 * g("k4pend") + 5 * g("gem") + 4 * g("k23e")
            - 14 * g("k4wing") - 12 * g("w4") + 30 * g("k5e") - 60 * g("k5"))
    dp3 = (g("dp3") - 6 * g("book3") - 3 * g("k4pend") - 2 * g("gem")
           + 10 * g("k4wing") + 4 * g("w4") - 24 * g("k5e") + 60 * g("k5"))
    paw = g("paw") - 4 * g("pairs_t") + 12 * g("k4")
    claw = g("claw_sums") + 2 * g("pairs_t") - 4 * g("k4")
    return {2: fork, 3: 4 * star + 3 * claw, 5: 2 * bull, 6: 2 * dp3 + 2 * paw,
            "induced": {"k14": star, "fork": fork, "bull": bull,
                        "diamond_pendant": dp3, "paw": paw, "claw": claw}}


@pytest.mark.parametrize("seed", range(24))
def test_p2_p3_p5_p6_closed_forms(seed):
    n, edges, adj = _graph(seed)
    N = _first_principles(adj)
    D = _dag_decomposition(adj)
    # the device's decomposition gives the same subgraph counts
    assert D["w4"] == N["w4"] and D["k23"] == N["k23"]
    assert D["k23e"] == N["k23e"]
    assert D["tq"] - 4 * N["pairs_t"] == N["house"]
    assert D["ban"] - 2 * N["pairs_t"] == N["banner"]
    got = _patterns(N)
    rp, ci = csr_from_edges(n, edges)
    brute = oracle.brute(rp, ci)
    for k, v in got["induced"].items():
        assert v == brute[k], k
    for p in (2, 3, 5, 6):
        assert got[p] == oracle.pattern(rp, ci, p)[0], p


def test_planted_mistake_fails():
    """Dropping the W4 term (the count the anchor pass exists for) breaks
    P3 on a graph that holds wheels: the pin detects it."""
    n, edges, adj = 9, [], [set() for _ in range(9)]
    for a, b in itertools.combinations(range(9), 2):
        if (a * 5 + b * 3) % 4 != 0:
            adj[a].add(b)
            adj[b].add(a)
            edges.append((a, b))
    N = _first_principles(adj)
    assert N["w4"] > 0
    rp, ci = csr_from_edges(n, edges)
    want = oracle.pattern(rp, ci, 3)[0]
    assert _patterns(N)[3] == want
    N["w4"] = 0
    assert _patterns(N)[3] != want


def test_containment_rows():
    """The coefficient rows in `_patterns` are rows of the inverse of the
    containment matrix c(H, H') (labelled copies of H inside H') of the 21
    connected graphs on five vertices, built here by brute force."""
    V = range(5)
    pairs = list(itertools.combinations(V, 2))
    perms = list(itertools.permutations(V))

    def canon(es):
        return min(tuple(sorted(tuple(sorted((p[a], p[b]))) for a, b in es))
                   for p in perms)

    def connected(es):
        adj = {i: set() for i in V}
        for a, b in es:
            adj[a].add(b)
            adj[b].add(a)
        seen, st = {0}, [0]
        while st:
            for y in adj[st.pop()]:
                if y not in seen:
                    seen.add(y)
                    st.append(y)
        return len(seen) == 5

    classes = {}
    for mask in range(1 << 10):
        es = [pairs[i] for i in range(10) if mask >> i & 1]
        if connected(es):
            classes.setdefault(canon(es), []).append(frozenset(es))
    keys = sorted(classes, key=lambda c: (len(c), c))
    assert len(keys) == 21
    Cm = np.array([[sum(1 for e in classes[h] if e <= set(hp)) for hp in keys]
                   for h in keys], dtype=np.float64)
    inv = np.rint(np.linalg.inv(Cm)).astype(np.int64)

    def key(edges):
        return keys.index(canon(edges))

    star = key([(0, 1), (0, 2), (0, 3), (0, 4)])
    fork = key([(0, 1), (0, 2), (0, 3), (1, 4)])
    bull = key([(0, 1), (0, 2), (0, 3), (1, 2), (1, 4)])
    dp3 = key([(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3)])
    name = {
        key([(0, 1), (0, 2), (0, 3), (0, 4)]): "star",
        key([(0, 1), (0, 2), (0, 3), (1, 4)]): "fork",
        key([(0, 1), (0, 2), (0, 3), (0, 4), (1, 2)]): "cricket",
        key([(0, 1), (0, 2), (0, 3), (1, 2), (1, 4)]): "bull",
        key([(0, 1), (0, 2), (0, 3), (1, 2), (3, 4)]): "lolli",
        key([(0, 1), (0, 2), (0, 3), (1, 4), (2, 4)]): "banner",
        key([(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3)]): "dp3",
        key([(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (3, 4)]): "bowtie",
        key([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 4)]): "dp2",
        key([(0, 1), (0, 2), (0, 3), (1, 2), (1, 4), (3, 4)]): "house",
        key([(0, 1), (0, 2), (0, 3), (1, 4), (2, 4), (3, 4)]): "k23",
        key([(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3), (1, 4)]): "book3",
        key([(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3), (2, 3)]):
            "k4pend",
        key([(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3), (2, 4)]): "gem",
        key([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 4), (3, 4)]): "k23e",
        key([(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3), (1, 4),
             (2, 3)]): "k4wing",
        key([(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3), (2, 4),
             (3, 4)]): "w4",
        key([(a, b) for a, b in pairs if (a, b) != (3, 4)]): "k5e",
        key(pairs): "k5",
    }
    for row, probe in ((star, "star"), (fork, "fork"), (bull, "bull"),
                       (dp3, "dp3")):
        N = {v: 0 for v in name.values()}
        for j in range(21):
            if inv[row, j]:
                N[name[j]] = int(inv[row, j])
        # _patterns is linear in N: unit vectors read its coefficients back
        got = {}
        for h in name.values():
            unit = {k: (1 if k == h else 0) for k in name.values()}
            unit.update(paw=0, pairs_t=0, k4=0, claw_sums=0)
            got[h] = _patterns(unit)["induced"][{
                "star": "k14", "fork": "fork", "bull": "bull",
                "dp3": "diamond_pendant"}[probe]]
        assert got == N, probe
