"""Device tree (include/sphk_dtree.h) against the host tree (sphk_tree.h).

Both builders implement the same octree, walks and list policies with the
same evaluation order, so on the same Morton-ordered input they must agree
element for element: walk table, node count, FP64 node monopoles (bitwise),
and the gather / symmetry / long lists (j and SPJ indices).  The host tree
itself is pinned against brute force in test_abi_tree.py.
"""
import numpy as np
import pytest
import torch

from research_codes_amd import _lib as L
from research_codes_amd import ic
from research_codes_amd.device import FP64, Context
from research_codes_amd.step import SPHSystem
from research_codes_amd.tree import DeviceTree, HostTree, morton_order, morton_order_dev, sort_particles

pytestmark = pytest.mark.gpu


def _cases():
    star = ic.polytrope(1.5, 3000, M=1.0, R=1.0, seed=11, eta=1.6)
    pair = ic.head_on(2500, eta=1.18)
    return {"star": star, "pair": pair}


@pytest.fixture(scope="module", params=["star", "pair"])
def case(request):
    p, _ = sort_particles(_cases()[request.param])
    n = len(p["x"])
    rng = np.random.default_rng(5)
    p["eps2"] = rng.uniform(1e-5, 1e-3, n)
    p["rs"] = 1.1 * p["ksr"] * rng.uniform(0.9, 1.1, n)
    return p


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_morton_order_matches_host():
    p = _cases()["pair"]
    perm_h = morton_order(p["x"], p["y"], p["z"])
    perm_d = morton_order_dev(_d(p["x"]), _d(p["y"]), _d(p["z"]))
    assert np.array_equal(perm_h, perm_d.cpu().numpy())


def test_unsorted_input_is_rejected():
    p = _cases()["star"]
    with pytest.raises(L.SphkError):
        DeviceTree(_d(p["x"]), _d(p["y"]), _d(p["z"]), _d(p["m"]))


@pytest.mark.parametrize("n_leaf,n_group", [(8, 16), (8, 64), (1, 8), (16, 256)])
def test_tree_and_lists_match_host(case, n_leaf, n_group):
    p = case
    print(f"[dtree] n={len(p['x'])} n_leaf={n_leaf} n_group={n_group}", flush=True)
    ht = HostTree(p["x"], p["y"], p["z"], p["m"], p["eps2"], n_leaf, n_group)
    dt = DeviceTree(_d(p["x"]), _d(p["y"]), _d(p["z"]), _d(p["m"]), _d(p["eps2"]), n_leaf, n_group)
    try:
        assert dt.n_walk == ht.n_walk and dt.n_node == ht.n_node
        assert np.array_equal(dt.i_off.cpu().numpy(), ht.i_off)
        hm, dm = ht.moments(), dt.moments()
        for k in ("m", "x", "y", "z", "eps2"):
            assert np.array_equal(dm[k].cpu().numpy(), hm[k]), k
        for name, hl, dl in (("gather", ht.gather(p["rs"]), dt.gather(_d(p["rs"]))),
                             ("symmetry", ht.symmetry(p["ksr"]), dt.symmetry(_d(p["ksr"]))),
                             ("long0.5", ht.long(0.5), dt.long(0.5)),
                             ("long0.3", ht.long(0.3), dt.long(0.3))):
            assert len(hl) == len(dl), name
            for a, b in zip(hl, dl):
                assert np.array_equal(a, b.cpu().numpy()), name
    finally:
        ht.close()
        dt.close()


def _fractal(levels=14):
    """Particles whose octree opens 8 children at every one of `levels`
    levels: 7 single-particle octants plus one that nests again; anchors at
    the corners of the unit cube."""
    pts = [(0.0, 0.0, 0.0), (1.0, 1.0, 1.0)]
    for k in range(1, levels + 1):
        h = 2.0 ** -k
        for a in range(1, 8):
            pts.append((((a >> 2) & 1) * h + 0.5 * h, ((a >> 1) & 1) * h + 0.5 * h, (a & 1) * h + 0.5 * h))
    x, y, z = (np.array(v) for v in zip(*pts))
    perm = morton_order(x, y, z)
    n = len(x)
    return {"x": x[perm], "y": y[perm], "z": z[perm], "m": np.full(n, 1.0 / n), "eps2": np.full(n, 1e-8),
            "h": np.full(n, 3.0)}


def test_deep_tree_lists_overflowing_the_small_stack():
    """Every walk of this tree descends 14 levels with 7 siblings pending at
    each (a preorder stack of about 100 entries): the device list walks
    overflow their small LDS stack and run again with the worst-case one; the
    lists must still equal the host's."""
    p = _fractal()
    ht = HostTree(p["x"], p["y"], p["z"], p["m"], p["eps2"], 1, 8)
    dt = DeviceTree(_d(p["x"]), _d(p["y"]), _d(p["z"]), _d(p["m"]), _d(p["eps2"]), 1, 8)
    try:
        assert dt.n_node == ht.n_node and dt.n_walk == ht.n_walk
        for name, hl, dl in (("gather", ht.gather(p["h"]), dt.gather(_d(p["h"]))),
                             ("symmetry", ht.symmetry(p["h"]), dt.symmetry(_d(p["h"]))),
                             ("long0", ht.long(0.0), dt.long(0.0))):
            assert len(hl) == len(dl), name
            for a, b in zip(hl, dl):
                assert np.array_equal(a, b.cpu().numpy()), name
    finally:
        ht.close()
        dt.close()


def test_walks_beyond_the_staged_i_particles():
    """40 coincident particles share one leaf at the deepest level, a walk of
    40 > max(n_group, n_leaf) = 8 particles: the member test stages 8 of a
    walk's i-particles and reads the rest from memory; the lists must still
    equal the host's."""
    rng = np.random.default_rng(3)
    n = 600
    x, y, z = rng.uniform(0.0, 1.0, (3, n))
    x[:40], y[:40], z[:40] = 0.37, 0.41, 0.53
    perm = morton_order(x, y, z)
    x, y, z = x[perm], y[perm], z[perm]
    m, eps2, h = np.full(n, 1.0 / n), np.full(n, 1e-6), rng.uniform(0.1, 0.2, n)
    ht = HostTree(x, y, z, m, eps2, 1, 8)
    dt = DeviceTree(_d(x), _d(y), _d(z), _d(m), _d(eps2), 1, 8)
    try:
        assert np.diff(ht.i_off).max() >= 40
        assert dt.n_walk == ht.n_walk and np.array_equal(dt.i_off.cpu().numpy(), ht.i_off)
        for name, hl, dl in (("gather", ht.gather(h), dt.gather(_d(h))),
                             ("symmetry", ht.symmetry(h), dt.symmetry(_d(h))),
                             ("long0.5", ht.long(0.5), dt.long(0.5))):
            assert len(hl) == len(dl), name
            for a, b in zip(hl, dl):
                assert np.array_equal(a, b.cpu().numpy()), name
    finally:
        ht.close()
        dt.close()


def test_rebuild_in_place_matches_fresh_build():
    """One handle rebuilt for a sequence of particle sets (larger, smaller,
    another leaf and group size) gives exactly what a new handle gives."""
    cases = _cases()
    seq = [(sort_particles(cases["star"])[0], 8, 32), (sort_particles(cases["pair"])[0], 8, 16),
           (sort_particles(cases["star"])[0], 4, 64)]
    seq.append((sort_particles({k: v[:1200] for k, v in cases["pair"].items()})[0], 8, 32))
    keep = None
    for p, n_leaf, n_group in seq:
        args = [_d(p[k]) for k in ("x", "y", "z", "m")] + [_d(p["ksr"] ** 2), n_leaf, n_group]
        print(f"[dtree] rebuild n={len(p['x'])} n_leaf={n_leaf} n_group={n_group}", flush=True)
        if keep is None:
            keep = DeviceTree(*args)
        else:
            keep.rebuild(*args)
        fresh = DeviceTree(*args)
        try:
            assert keep.n_walk == fresh.n_walk and keep.n_node == fresh.n_node
            assert torch.equal(keep.i_off, fresh.i_off)
            for k in ("m", "x", "y", "z", "eps2"):
                assert torch.equal(keep.moments()[k], fresh.moments()[k]), k
            rs = _d(1.1 * p["ksr"])
            for name, a, b in (("gather", keep.gather(rs), fresh.gather(rs)),
                               ("symmetry", keep.symmetry(_d(p["ksr"])), fresh.symmetry(_d(p["ksr"]))),
                               ("long", keep.long(0.5), fresh.long(0.5))):
                for u, v in zip(a, b):
                    assert torch.equal(u, v), name
        finally:
            fresh.close()
    keep.close()


def test_step_with_device_lists_equals_host_lists():
    """The same lists give bitwise the same force step."""
    p = ic.head_on(3000, eta=1.18)
    out = []
    for tree in ("host", "device"):
        with Context(0) as c:
            s = SPHSystem(p, c, eta=1.18, eps=0.01, precision=FP64, n_group=16, tree=tree)
            s.initialize()
            out.append({k: s.s[k].clone() for k in ("dens", "ksr", "ax", "ay", "az", "udot", "pot")})
    for k in out[0]:
        assert torch.equal(out[0][k], out[1][k]), k


def test_device_resort_matches_host():
    p = ic.head_on(2000, eta=1.18)
    res = []
    for tree in ("host", "device"):
        with Context(0) as c:
            s = SPHSystem(p, c, eta=1.18, eps=0.01, precision=FP64, n_group=16, tree=tree)
            g = torch.Generator(device="cpu").manual_seed(3)
            shift = (torch.rand(s.n, generator=g, dtype=torch.float64) - 0.5).cuda() * 0.05
            s.s["x"].add_(shift)
            s.rs.copy_(torch.arange(s.n, dtype=torch.float64, device="cuda"))
            s.np_.copy_(torch.arange(s.n, device="cuda") * 3)
            s.itr.copy_((torch.arange(s.n, device="cuda") % 7).to(torch.uint8))
            s.eps2.copy_(s.s["y"] * 2.0)
            s.resort()
            assert s.last_perm is not None
            res.append([s.id, s.istar, s._sbuf, s.rs, s.np_, s.itr, s.eps2])
    for a, b in zip(*res):
        assert torch.equal(a, b)


@pytest.mark.parametrize("n", [1, 257, 10001])
def test_permute_and_equal_entry_points(n):
    """sphk_permute_dev against torch indexing on columns of 1, 2, 4 and 8
    bytes (rows of a 2D array with an odd length are only 8-byte aligned);
    sphk_equal_dev is bitwise: -0.0 differs from 0.0, a NaN equals itself."""
    from research_codes_amd.tree import permute_dev
    g = torch.Generator(device="cpu").manual_seed(n)
    perm = torch.randperm(n, generator=g).to(torch.int32).cuda()
    rows = torch.rand(3, n, generator=g, dtype=torch.float64).cuda()
    cols = [rows[0], rows[1], rows[2], torch.randint(0, 255, (n,), generator=g, dtype=torch.uint8).cuda(),
            torch.randint(-9, 9, (n,), generator=g, dtype=torch.int16).cuda(),
            torch.randint(-9, 9, (n,), generator=g, dtype=torch.int32).cuda(),
            torch.randint(-9, 9, (n,), generator=g, dtype=torch.int64).cuda()]
    want = [c[perm.long()].clone() for c in cols]
    work = permute_dev(perm, cols)
    work = permute_dev(torch.arange(n, dtype=torch.int32, device="cuda"), cols, work)   # identity, reused scratch
    for c, w in zip(cols, want):
        assert torch.equal(c, w)
    a = torch.rand(n, generator=g, dtype=torch.float64).cuda()
    flag = torch.ones(1, dtype=torch.int32, device="cuda")

    def equal(x, y, count=2):
        A = (L._P * count)(*[x.data_ptr()] * count)
        B = (L._P * count)(*[y.data_ptr()] * count)
        L.check(L.lib().sphk_equal_dev(n, count, A, B, flag.data_ptr(), torch.cuda.current_stream().cuda_stream),
                "sphk_equal_dev")
        return int(flag.item()) == 0

    b = a.clone()
    assert equal(a, b) and equal(a, b, 0)
    b[-1] = float("nan")
    a[-1] = float("nan")
    assert equal(a, b)
    a[-1], b[-1] = 0.0, -0.0
    assert not equal(a, b)
    b[-1] = 0.0
    b[0] = torch.nextafter(b[0], torch.tensor(2.0, dtype=torch.float64, device="cuda"))
    assert not equal(a, b)


def test_short_lists_beyond_the_walk_room():
    """The short-range lists are built in one walk into a fixed room of
    candidates per walk; a build whose walks overflow it runs the counted
    build instead, and later builds get a larger room.  Wide search radii
    and groups of 256 overflow the first room: every build, before and after
    the room grows, must still equal the host's lists."""
    p, _ = sort_particles(_cases()["star"])
    ht = HostTree(p["x"], p["y"], p["z"], p["m"], None, 8, 256)
    dt = DeviceTree(_d(p["x"]), _d(p["y"]), _d(p["z"]), _d(p["m"]), None, 8, 256)
    try:
        for fac in (4.0, 4.0, 1.1, 6.0):
            rs = fac * p["ksr"]
            hl, dl = ht.gather(rs), dt.gather(_d(rs))
            n_cand_max = int(np.diff(hl[0]).max())
            print(f"[dtree room] fac {fac}: longest gather list {n_cand_max}", flush=True)
            for a, b in zip(hl, dl):
                assert np.array_equal(a, b.cpu().numpy()), fac
            hs, ds = ht.symmetry(rs), dt.symmetry(_d(rs))
            for a, b in zip(hs, ds):
                assert np.array_equal(a, b.cpu().numpy()), fac
    finally:
        ht.close()
        dt.close()


def test_long_lists_beyond_the_walk_room():
    """The long list is built in one walk into rooms of particle (EPJ) and
    node (SPJ) entries per walk, then packed.  Direct summation (theta 0) of
    10 000 particles overflows the first EPJ room (8192): the counted build
    runs instead and later builds get twice the room; every build must equal
    the host's lists."""
    p = ic.polytrope(1.5, 10000, M=1.0, R=1.0, seed=12, eta=1.6)
    p, _ = sort_particles(p)
    ht = HostTree(p["x"], p["y"], p["z"], p["m"], None, 8, 64)
    dt = DeviceTree(_d(p["x"]), _d(p["y"]), _d(p["z"]), _d(p["m"]), None, 8, 64)
    try:
        for theta in (0.0, 0.0, 0.5, 0.3):
            hl, dl = ht.long(theta), dt.long(theta)
            print(f"[dtree room] theta {theta}: longest EPJ list {int(np.diff(hl[0]).max())}", flush=True)
            assert len(hl) == len(dl)
            for a, b in zip(hl, dl):
                assert np.array_equal(a, b.cpu().numpy()), theta
    finally:
        ht.close()
        dt.close()


def test_handle_arrays_of_lists_built_in_one_walk():
    """A list built in one walk stays in its rooms until read: the handle's
    own arrays (sphk_dtree_list) pack it, also after another list has been
    built on the same handle (the rooms are reused), and equal the copy."""
    import ctypes as C
    p = sort_particles(_cases()["pair"])[0]
    dt = DeviceTree(_d(p["x"]), _d(p["y"]), _d(p["z"]), _d(p["m"]), None, 8, 32)
    try:
        g = dt.gather(_d(1.1 * p["ksr"]))
        lg = dt.long(0.5)

        def handle_arrays(kind):
            jo, ji, so, si = (C.c_void_p() for _ in range(4))
            nj, ns = C.c_int64(), C.c_int64()
            assert L.lib().sphk_dtree_list(dt._h, kind, C.byref(jo), C.byref(ji), C.byref(so), C.byref(si),
                                           C.byref(nj), C.byref(ns)) == 0
            out = [(jo.value, dt.n_walk + 1), (ji.value, nj.value)]
            if kind == 2:
                out += [(so.value, dt.n_walk + 1), (si.value, ns.value)]
            res = []
            for ptr, n in out:
                t = torch.empty(n, dtype=torch.int32, device="cuda")
                assert L.lib().sphk_copy_dev(t.data_ptr(), ptr, 4 * n, None) == 0
                res.append(t)
            torch.cuda.synchronize()
            return res

        for a, b in zip(handle_arrays(0), g):
            assert torch.equal(a, b)
        for a, b in zip(handle_arrays(2), lg):
            assert torch.equal(a, b)
        # read again after the handle packed them: the copy path
        for a, b in zip(dt._get(2), lg):
            assert torch.equal(a, b)
    finally:
        dt.close()
