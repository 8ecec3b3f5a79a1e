"""Model check of the local union that k_merge runs in LDS per wave (csrc/swmi_kernels.hip, MergeEdge).

Each lane unites the two local nodes of one root pair: it chases both to their roots, links the root with the
larger key under the other with one atomicMin on the packed entry (key of the parent << 32 | parent), and when
the atomicMin returns an entry other than the root entry it read, it continues by uniting its own root with the
parent the atomicMin returned (the link it may have displaced). The model runs the lanes' steps (chase, then
atomicMin) in a random interleaving, so chase results may be stale when the atomic executes, and checks that
the resulting forest has the connected components of the pairs, with every parent carrying a smaller key. The
same model with the continuation left out loses links, which the last test shows.
"""

import random

import pytest


def _unite(pairs, seed, continue_with_displaced=True):
    rng = random.Random(seed)
    nodes = sorted({x for p in pairs for x in p})
    key = {x: rng.randrange(1 << 20) * 1000 + i for i, x in enumerate(nodes)}   # distinct keys, as sites are
    entry = {x: (key[x], x) for x in nodes}                                     # a root holds its own key
    lanes = [{"a": a, "b": b, "open": True, "chased": False} for a, b in pairs]
    linked = []                                                                 # successful links (b under a)

    def root(x):
        while entry[x][1] != x:
            x = entry[x][1]
        return x

    for _ in range(200000):
        live = [ln for ln in lanes if ln["open"]]
        if not live:
            break
        ln = rng.choice(live)
        if not ln["chased"]:
            ra, rb = root(ln["a"]), root(ln["b"])
            ln.update(ra=ra, rb=rb, ea=entry[ra], eb=entry[rb])
            if ra == rb:
                ln["open"] = False
            else:
                ln["chased"] = True
            continue
        a, b, ea, eb = ln["ra"], ln["rb"], ln["ea"], ln["eb"]
        if ea[0] > eb[0]:
            a, b, ea, eb = b, a, eb, ea
        old = entry[b]
        entry[b] = min(old, (ea[0], a))     # atomicMin on the packed entry
        ln["chased"] = False
        if old == eb:
            ln["open"] = False
            linked.append((b, a))
        elif continue_with_displaced:
            ln["a"], ln["b"] = a, old[1]
        else:
            ln["a"], ln["b"] = a, b
    else:
        raise AssertionError("lanes did not finish")
    return nodes, key, entry, root, linked


def _components(pairs, nodes):
    par = {x: x for x in nodes}

    def find(x):
        while par[x] != x:
            x = par[x]
        return x

    for a, b in pairs:
        par[find(a)] = find(b)
    return find


def _check(pairs, seed, continue_with_displaced=True):
    nodes, key, entry, root, _ = _unite(pairs, seed, continue_with_displaced)
    find = _components(pairs, nodes)
    for x in nodes:
        if entry[x][1] != x:
            assert key[entry[x][1]] < key[x] and entry[x][0] == key[entry[x][1]]
    return all((find(a) == find(b)) == (root(a) == root(b)) for a in nodes for b in nodes)


def _pair_sets(count):
    for t in range(count):
        rng = random.Random(t)
        n = rng.randint(2, 16)
        yield t, [tuple(rng.sample(range(n), 2)) for _ in range(rng.randint(1, 40))]


def test_local_union_keeps_components_under_any_interleaving():
    for t, pairs in _pair_sets(1500):
        assert _check(pairs, t), pairs


def test_local_union_star_around_one_slot():
    # one slot paired with many others (a cluster of one block meeting many of its neighbour's): lanes race on
    # one root entry whenever that slot's key is not the smallest
    for t in range(200):
        pairs = [(0, j) for j in range(1, 30)]
        assert _check(pairs, t)


@pytest.mark.parametrize("count", [1500])
def test_model_detects_a_union_that_drops_the_displaced_parent(count):
    # the retry without the returned parent loses links in some interleaving
    assert any(not _check(pairs, t, continue_with_displaced=False) for t, pairs in _pair_sets(count))


def test_successful_local_links_span_the_components():
    # each successful link joins two different trees, so the links alone connect every component of the pairs
    # (the alternative of making each lane's own link lattice-wide relies on this)
    for t, pairs in _pair_sets(1500):
        nodes, _, _, _, linked = _unite(pairs, t)
        find_pairs = _components(pairs, nodes)
        find_links = _components(linked, nodes)
        assert len(linked) == len(nodes) - len({find_pairs(x) for x in nodes})
        assert all((find_pairs(a) == find_pairs(b)) == (find_links(a) == find_links(b)) for a in nodes for b in nodes)
