"""The store's block format on the GPU (octree.h:35-43, octree.cpp:427-559)
against the oracle: blocks of a LAS cloud's tree (records' fields, merged
locations, classes) and of a tree without records, byte for byte; the
store's files; blocks read back and rebuilt into the same tree."""
import os

import numpy as np
import pytest

import las_spec
import oracle_lib as O
import oracle_las as OL
import wolkenbase_amd as W
from wolkenbase_amd import las as WL
from wolkenbase_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = W.Context(0)
    yield c
    c.close()


def las_with_duplicates(seed=71):
    """A format-3 tile with repeated locations carrying other attributes."""
    rng = np.random.default_rng(13)
    cloud = synth.generate(seed=seed, extent=20.0, density=30.0)
    n = len(cloud["xi"])
    f = las_spec.random_fields(rng, n, 3, xyz_range=10)
    for k, r in enumerate(f):
        r["xi"], r["yi"], r["zi"] = int(cloud["xi"][k]), int(cloud["yi"][k]), int(cloud["zi"][k])
        r["ret"] = 1 + r["ret"] % 3
    for k in rng.integers(0, n, n // 10):
        g = dict(f[int(k)])
        g["intensity"] = int(rng.integers(0, 65536))
        g["red"] = int(rng.integers(0, 65536))
        f.insert(int(rng.integers(0, len(f))), g)
    return las_spec.image(3, f, scale=tuple(cloud["scale"]), offset=tuple(cloud["offset"]),
                          bounds=(tuple(cloud["bbmin"]), tuple(cloud["bbmax"])))


@pytest.mark.parametrize("classify", [True, False])
def test_blocks_of_a_cloud_match_oracle(ctx, classify):
    img = las_with_duplicates()
    files = [WL.LasFile(ctx, image=img)]
    cl = WL.Cloud(ctx, files)
    tree = cl.build()
    cls = cl.classify(tree, want_tiles=False)[0] if classify else None
    got = tree.store_blocks(cloud=cl, cls=cls)
    o = OL.pipeline([img], day=1, year=2020, classify=classify)
    assert tree.ndup == o["ndup"] > 0
    want = OL.store_blocks(o["stored"], o["leaves"]["count"])
    assert len(got) == 32768 * len(o["leaves"]["count"])
    assert got.tobytes() == want
    tree.free()
    cl.close()


def test_blocks_without_records_match_oracle(ctx):
    cloud = synth.generate(seed=5, extent=30.0, density=25.0, outliers=0.01)
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    tree = W.Octree.build(ctx, cloud["xi"], cloud["yi"], cloud["zi"], cloud["scale"],
                          cloud["offset"], center, side)
    cls, _, _ = W.classify(tree, cloud["bbmin"], cloud["bbmax"])
    got = tree.store_blocks(cls=cls)
    x = cloud["offset"][0] + cloud["scale"][0] * cloud["xi"].astype(np.float64)
    y = cloud["offset"][1] + cloud["scale"][1] * cloud["yi"].astype(np.float64)
    z = cloud["offset"][2] + cloud["scale"][2] * cloud["zi"].astype(np.float64)
    ref = O.Tree(x, y, z, center, side)
    lv = ref.leaves()
    order = ref.order()
    stored = []
    for i in order:
        p = OL.Point()
        p.x, p.y, p.z = float(x[i]), float(y[i]), float(z[i])
        p.cls = int(cls[i])
        stored.append(p)
    assert got.tobytes() == OL.store_blocks(stored, lv["count"])
    tree.free()


def test_files_and_read_back(ctx, tmp_path):
    img = las_with_duplicates(seed=72)
    cl = WL.Cloud(ctx, [WL.LasFile(ctx, image=img)])
    tree = cl.build()
    cls = cl.classify(tree, want_tiles=False)[0]
    blocks = tree.store_blocks(cloud=cl, cls=cls)
    nl = tree.info()["nleaves"]
    # the store's files: block b in file b % 3 at 32768 * (b // 3)
    name = os.path.join(str(tmp_path), "store")
    tree.store_write_files(name, 3, cloud=cl, cls=cls)
    for f in range(3):
        data = open(name + str(f), "rb").read()
        own = [b for b in range(nl) if b % 3 == f]
        assert len(data) == 32768 * len(own)
        for j, b in enumerate(own):
            assert data[32768 * j:32768 * (j + 1)] == blocks[32768 * b:32768 * (b + 1)].tobytes()
    # OctBuffer::read of every block: the stored points of each leaf
    x, y, z, c, counts = ctx.store_read_blocks(blocks)
    lv = tree.leaves()
    assert np.array_equal(counts, lv["count"])
    o = OL.pipeline([img], day=1, year=2020)
    sx = np.array([p.x for p in o["stored"]])
    sy = np.array([p.y for p in o["stored"]])
    sz = np.array([p.z for p in o["stored"]])
    sc = np.array([p.cls for p in o["stored"]], dtype=np.uint16)
    assert np.array_equal(x, sx) and np.array_equal(y, sy) and np.array_equal(z, sz)
    assert np.array_equal(c, sc)
    # the points read back, in block order, build the same tree, its
    # canonical order being the order read
    center, side = cl.size_fit()
    t2 = W.Octree.build_xyz(ctx, x, y, z, o["resolution"], center, side)
    l2 = t2.leaves()
    for k in ("count", "side", "depth"):
        assert np.array_equal(l2[k], lv[k]), k
    assert np.array_equal(l2["center"], lv["center"])
    assert np.array_equal(t2.order(), np.arange(len(x), dtype=np.uint32))
    # a block never written reads as empty
    x0, _, _, _, k0 = ctx.store_read_blocks(bytes(32768 * 2))
    assert len(x0) == 0 and list(k0) == [0, 0]
    t2.free()
    tree.free()
    cl.close()
