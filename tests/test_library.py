"""The workload generators on the host: the random benchmark circuit's
no-repeat rule and gate counts (24/30/32 qubits at depth 20: 710/890/950),
QFT sizes, the Eagle heavy-hex coupling map (library.rs:388-425: 127
qubits, 144 couplers) and its 33-qubit restriction (35 couplers), and the
kicked-Ising circuit built on it (library.rs:430-446)."""
from collections import Counter, deque

import pytest

from qevedo.simulator import library

NAMES = {"rx": 0, "ry": 1, "u3": 2}


def layers_of(gates, n):
    """Split the random circuit into its layers: n one-qubit gates then the
    CZ brick of that layer."""
    out, cur = [], []
    for g in gates:
        if g[0] != "cz" and cur and cur[-1][0] == "cz":
            out.append(cur)
            cur = []
        cur.append(g)
    out.append(cur)
    return out


@pytest.mark.parametrize("n,seed", [(5, 1), (24, 42), (30, 42), (32, 7)])
def test_random_circuit_structure_and_no_repeat(n, seed):
    depth = 20
    gates = library.random_circuit(n, depth, seed)
    layers = layers_of(gates, n)
    assert len(layers) == depth
    previous = [None] * n
    uses = Counter()
    for layer, lg in enumerate(layers):
        ones = [g for g in lg if g[0] != "cz"]
        czs = [g for g in lg if g[0] == "cz"]
        assert [g[1] for g in ones] == [(q,) for q in range(n)]
        assert czs == [("cz", (q, q + 1), ()) for q in range(layer % 2, n - 1, 2)]
        for q, (name, _, params) in enumerate(ones):
            assert (name, params) in library.SQRT_GATES
            g = NAMES[name]
            assert g != previous[q], (layer, q)  # never the qubit's previous gate
            previous[q] = g
            uses[g] += 1
    # all three gates appear, none dominates (drawn uniformly)
    assert len(uses) == 3 and max(uses.values()) < 0.5 * n * depth


def test_random_circuit_gate_counts():
    # n one-qubit gates per layer plus floor((n - layer % 2) / 2) CZs
    for n, want in [(24, 710), (30, 890), (32, 950)]:
        assert len(library.random_circuit(n, 20, 42)) == want
        assert want == 20 * n + sum((n - l % 2) // 2 for l in range(20))


def test_sqrt_w_is_u3_of_the_reference():
    # library.rs:41-50: sqrt(X) = RX(pi/2), sqrt(Y) = RY(pi/2),
    # sqrt(W) = U(pi/2, -pi/4, pi/4)
    import math
    assert library.SQRT_GATES[0] == ("rx", (math.pi / 2,))
    assert library.SQRT_GATES[1] == ("ry", (math.pi / 2,))
    assert library.SQRT_GATES[2] == ("u3", (math.pi / 2, -math.pi / 4, math.pi / 4))


def test_qft_sizes():
    for n in (1, 2, 9, 30):
        gates = library.qft(n)
        assert len(gates) == n * (n + 1) // 2
    assert len(library.qft(30)) == 465
    g = library.qft(3)
    assert g[0] == ("h", (2,), ())
    assert g[1][0] == "cp" and g[1][1] == (1, 2)


def _components(nodes, edges):
    adj = {v: set() for v in nodes}
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    seen, comps = set(), 0
    for v in nodes:
        if v in seen:
            continue
        comps += 1
        dq = deque([v])
        seen.add(v)
        while dq:
            u = dq.popleft()
            for w in adj[u]:
                if w not in seen:
                    seen.add(w)
                    dq.append(w)
    return comps


def test_eagle_heavy_hex():
    e = library.eagle_heavy_hex()
    nodes = sorted({q for ed in e for q in ed})
    assert nodes == list(range(127))
    assert len(e) == 144 and len(set(map(frozenset, e))) == 144
    deg = Counter(q for ed in e for q in ed)
    assert max(deg.values()) == 3
    assert _components(nodes, e) == 1
    # heavy hex is bipartite
    color = {0: 0}
    for _ in range(127):
        for a, b in e:
            if a in color and b not in color:
                color[b] = 1 - color[a]
            elif b in color and a not in color:
                color[a] = 1 - color[b]
    assert all(color[a] != color[b] for a, b in e)
    # the 24 bridge qubits (four between each pair of neighbouring rows) join
    # one qubit above and one below
    bridges = [q for q in nodes if deg[q] == 2 and all(abs(q - w) > 1 for ed in e if q in ed for w in ed if w != q)]
    assert len(bridges) == 24
    # couplers of the published ibm_kyiv / Eagle coupling map
    for pair in [(0, 14), (14, 18), (4, 15), (15, 22), (8, 16), (16, 26), (12, 17), (17, 30), (13, 12),
                 (32, 36), (36, 51), (109, 114), (96, 109), (125, 126)]:
        assert tuple(sorted(pair)) in {tuple(sorted(x)) for x in e}, pair


def test_heavy_hex_restricted_to_33_qubits():
    e = library.heavy_hex_edges(33)
    assert len(e) == 35
    assert all(a < 33 and b < 33 for a, b in e)
    assert _components(range(33), e) == 1
    assert max(Counter(q for ed in e for q in ed).values()) == 3


def test_kicked_ising_counts():
    e = library.heavy_hex_edges(33)
    gates = library.kicked_ising(33, e, 20, 0.7)
    assert len(gates) == 1360 == 20 * (33 + 35)
    assert gates[0] == ("rx", (0,), (0.7,))
    assert gates[33][0] == "rzz" and gates[33][2] == (-library.HALF_PI,)
