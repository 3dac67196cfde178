"""Surface reactions beyond uniform growth in the oracle (calcul.cpp:155-205):
oxidation with sphere removal (aggregat.cpp:230-243, 629-653;
aggregat_list.cpp:549-579), splits into connected parts (aggregat.cpp:553-628,
aggregat_list.cpp:413-444), reactions of the moved aggregate alone
(individual_surf_reactions) and partial updates between whole ones
(full_aggregate_update_frequency).  Pinned without the oracle's own code:

* under oxidation every sphere's radius is its previous radius plus u_sg times
  the step's time increment, and the spheres at or below rp_min leave the
  storage in order (numpy replay of the whole run, bit for bit);
* after every step of an oxidation run each aggregate's spheres form one
  connected set under the reference's contact test (periodic squared distance
  minus squared radius sum at most 1e-28), and each aggregate's volume is the
  list-order sum of its spheres' volumes;
* with individual reactions only the moved aggregate's spheres change, each by
  u_sg times its own time step scaled by the fraction of the move made;
* with partial updates an aggregate's volume stays at its last whole update
  until the next step whose count of steps without an event is a multiple of
  the frequency.
"""
import pathlib

import numpy as np
import pytest

import oracle_lib as O

CFG = pathlib.Path(__file__).parent / "configs"
DENSITY = 1800.0
RP_MIN = 0.166e-09
CONTACT_EPS = 1e-28


def reactions_ini(flux, n=300, seed=31, stop=1, extra=""):
    text = (CFG / "growth_small.ini").read_text()
    text = text.replace("flux_surfgrowth = 1e-3", f"flux_surfgrowth = {flux!r}")
    text = text.replace("number_of_aggregates = 200", f"number_of_aggregates = {stop}")
    text = text.replace("number = 2000", f"number = {n}").replace("random_seed = 31", f"random_seed = {seed}")
    return text + extra


def members_of(agg, a):
    return agg["members"][agg["offset"][a]:agg["offset"][a] + agg["n"][a]]


def connected(sph, mem, L):
    """One connected set under sphere_distances.cpp:81-90 (periodic distance)."""
    p = np.stack([sph["x"][mem], sph["y"][mem], sph["z"][mem]], axis=1)
    r = sph["r"][mem]
    d = p[:, None, :] - p[None, :, :]
    d = d - L * np.round(d / L)
    touch = (d ** 2).sum(axis=2) - (r[:, None] + r[None, :]) ** 2 <= CONTACT_EPS
    seen = np.zeros(len(mem), bool)
    stack = [0]
    seen[0] = True
    while stack:
        i = stack.pop()
        nxt = np.nonzero(touch[i] & ~seen)[0]
        seen[nxt] = True
        stack.extend(nxt.tolist())
    return bool(seen.all())


def list_sum(v):
    s = 0.0
    for x in v:
        s = s + x
    return s


def test_oxidation_shrinks_and_removes_spheres_in_order():
    flux = -30.0
    u_sg = flux / DENSITY
    sim = O.OracleSim(reactions_ini(flux, n=120, seed=5), fast_init=True)
    sph, _ = sim.export()
    r = sph["r"].copy()  # radii of the spheres still stored, in storage order
    removed = 0
    for step in range(4000):
        if sim.finished():
            break
        assert sim.run(1) == 1
        dt = sim.steps()[-1]["dt"]
        r = r + u_sg * dt
        keep = r > RP_MIN
        removed += int((~keep).sum())
        r = r[keep]
        got, agg = sim.export()
        assert np.array_equal(got["r"], r), f"step {step}"
        assert (agg["n"] > 0).all()
        assert np.array_equal(np.sort(agg["members"]), np.arange(len(r)))
    assert removed >= 60  # most of the monomers vanished
    assert sim.scalar("time") == pytest.approx(sim.steps()["dt"].sum(), rel=1e-12)


@pytest.mark.parametrize("flux,seed,steps", [(-3e-11, 32, 2500), (-3e-10, 32, 1500)])
def test_oxidation_splits_leave_connected_aggregates(flux, seed, steps):
    """Besides connectivity: the parts of a split are created in the order of
    their first sphere in the split aggregate's list, and each part starts
    with its earliest sphere of that list (a component grows from the first
    unvisited sphere)."""
    sim = O.OracleSim(reactions_ini(flux, n=300, seed=seed), fast_init=True)
    L = sim.scalar("box_length")
    sph0, agg0 = sim.export()
    splits = multi = 0
    for step in range(steps):
        assert sim.run(1) == 1
        sph, agg = sim.export()
        old_ids = set(agg0["uid"].tolist())
        new = [a for a in range(len(agg["n"])) if agg["uid"][a] not in old_ids]
        parents = {}
        for a in new:
            mem = members_of(agg, a)
            owner = [p for p in range(len(agg0["n"])) if np.isin(mem, members_of(agg0, p)).all()]
            if len(owner) != 1:
                continue  # formed by the step's merge
            plist = members_of(agg0, owner[0]).tolist()
            pos = [plist.index(m) for m in mem]
            assert pos[0] == min(pos), f"step {step}: a part does not start with its earliest sphere"
            parents.setdefault(owner[0], []).append((int(agg["uid"][a]), pos[0]))
        for parts in parents.values():
            if len(parts) < 2:
                continue  # the other parts merged in the same step
            assert [u for u, _ in sorted(parts, key=lambda t: t[1])] == sorted(u for u, _ in parts)
            splits += 1
        for a in range(len(agg["n"])):
            mem = members_of(agg, a)
            if len(mem) > 1:
                multi += 1
                assert connected(sph, mem, L), f"step {step} aggregate {agg['uid'][a]}"
            assert agg["volume"][a] == list_sum(sph["vol"][mem])
        sph0, agg0 = sph, agg
    assert splits >= 3 and multi > 0  # aggregates formed and split again


def test_merge_with_a_split_partner_is_refused_as_the_reference_does():
    """AggregatList::merge (aggregat_list.cpp:367-396) throws MergeError when an
    aggregate of the contact has split in the same step while its spheres
    still touch: the run ends with an error."""
    sim = O.OracleSim(reactions_ini(-1e-11, n=300, seed=31), fast_init=True)
    with pytest.raises(RuntimeError):
        sim.run(30000)
    assert 5000 < sim.scalar("steps") < 30000


def test_individual_reactions_change_the_moved_aggregate_only():
    flux = 1e-3
    u_sg = flux / DENSITY
    sim = O.OracleSim(reactions_ini(flux, n=300, seed=9, extra="[numerics]\nindividual_surf_reactions = true\n"),
                      fast_init=True)
    checked = 0
    for _ in range(400):
        sph0, agg0 = sim.export()
        assert sim.run(1) == 1
        st = sim.steps()[-1]
        sph1, agg1 = sim.export()
        if st["merged"]:
            continue
        mover = int(np.nonzero(agg0["uid"] == st["agg"])[0][0])
        mem = members_of(agg0, mover)
        factor = (st["dist"] if st["contact"] else st["lpm"]) / st["lpm"] + 0.0
        dt_indiv = agg0["time_step"][mover] * factor
        want = sph0["r"].copy()
        want[mem] = want[mem] + u_sg * dt_indiv
        assert np.array_equal(sph1["r"], want)
        others = np.ones(len(agg0["n"]), bool)
        others[mover] = False
        assert np.array_equal(agg1["volume"][others], agg0["volume"][others])
        assert agg1["volume"][mover] == list_sum(sph1["vol"][mem])
        checked += 1
    assert checked > 300


def test_partial_updates_keep_the_volume_until_the_next_whole_update():
    text = reactions_ini(1e-3, n=300, seed=13).replace("volsurf_method = none",
                                                       "volsurf_method = none\nfull_aggregate_update_frequency = 3")
    sim = O.OracleSim(text, fast_init=True)
    whole = partial = 0
    for _ in range(300):
        k = int(sim.scalar("n_iter_without_event"))
        sph0, agg0 = sim.export()
        assert sim.run(1) == 1
        st = sim.steps()[-1]
        sph1, agg1 = sim.export()
        if st["merged"]:
            continue
        assert not np.array_equal(sph1["r"], sph0["r"])  # every sphere grew
        sums = np.array([list_sum(sph1["vol"][members_of(agg1, a)]) for a in range(len(agg1["n"]))])
        if k % 3 == 0:
            assert np.array_equal(agg1["volume"], sums)
            whole += 1
        else:
            assert np.array_equal(agg1["volume"], agg0["volume"])
            assert not np.array_equal(agg1["volume"], sums)
            partial += 1
    assert whole > 50 and partial > 100


def test_split_components_follow_the_reference_stack():
    """Aggregate::split (aggregat.cpp:553-577) grows each part from its first
    unvisited sphere with a stack: the sphere popped last scans the unvisited
    ones in index order.  On this tree (0-1, 0-3, 3-2, 1-4, and 5 alone) that
    order is 0, 1, 3, 2, 4, worked by hand: 0 finds 1 and 3; 3, popped first,
    finds 2; 2 finds nothing; 1 finds 4.  A queue would give 0, 1, 3, 4, 2 and a
    recursive search 0, 1, 4, 3, 2."""
    pos = np.array([[0, 0, 0], [1.9, 0, 0], [0, 3.8, 0], [0, 1.9, 0], [3.8, 0, 0], [50, 50, 50]], float)
    r = np.ones(6)
    assert O.contact_components(pos, r, 1000.0) == [[0, 1, 3, 2, 4], [5]]
    # a part need not hold sphere 0's neighbours first: 0 alone, then 1-2
    pos2 = np.array([[0, 0, 0], [10, 0, 0], [11.9, 0, 0]], float)
    assert O.contact_components(pos2, np.ones(3), 1000.0) == [[0], [1, 2]]


def test_split_components_are_the_contact_graph_components():
    """On random clusters: the parts are the connected components of the
    contact graph (union-find over the periodic contact test,
    sphere_distances.cpp:81-90), ordered by their first sphere, each starting
    with its smallest index, and every later sphere touches an earlier one of
    its part."""
    rng = np.random.default_rng(7)
    L = 9.0
    for trial in range(300):
        n = int(rng.integers(2, 13))
        pos = rng.uniform(0, L, (n, 3))
        r = rng.uniform(0.6, 1.4, n)
        d = pos[:, None, :] - pos[None, :, :]
        d = d - L * np.round(d / L)
        touch = (d ** 2).sum(axis=2) - (r[:, None] + r[None, :]) ** 2 <= CONTACT_EPS
        parent = list(range(n))

        def find(i):
            while parent[i] != i:
                parent[i] = parent[parent[i]]
                i = parent[i]
            return i
        for i in range(n):
            for j in range(i + 1, n):
                if touch[i, j]:
                    parent[find(i)] = find(j)
        want = {}
        for i in range(n):
            want.setdefault(find(i), set()).add(i)
        got = O.contact_components(pos, r, L)
        assert sorted(map(frozenset, got), key=min) == sorted(map(frozenset, want.values()), key=min), trial
        assert [c[0] for c in got] == sorted(min(c) for c in want.values())
        for c in got:
            assert c[0] == min(c)
            for k in range(1, len(c)):
                assert touch[c[k], c[:k]].any(), (trial, c)
