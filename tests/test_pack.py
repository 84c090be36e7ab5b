"""SpherePack::makeCloud (lib/sphere-pack/SpherePack.cpp:135-248) on the device.

The oracle restates the reference's sequential algorithm (every candidate tested against
every placed sphere); it is pinned by closed forms and invariants: the radius sequence for
num > 0 ((num - i + 1/2)/(num + 1) through the mode's rule, decreasing), the mean radius
from the porosity, no two spheres overlapping (periodic: minimum image), every centre inside
its box (aperiodic: at least r from the walls), r^-2 affine in the quantile with distributeMass,
the solid volume V (1-p) that a mass-weighted PSD with its scaling fills, the generator's sequence (Park-Miller
x <- 48271 x mod 2^31-1: after 10000 draws from 1 the state is 399268537, the known value of
the minimal standard's 10000th output from seed 1), a PSD's radii within its range and the
retry at a higher porosity when the box cannot hold num spheres.  The device packing is
compared with the oracle's sphere for sphere, bit for bit; radii drawn by mass as the packing
goes (num < 0) to an ulp, since the C library's pow, which the oracle calls like the reference,
is not correctly rounded and the device's t^(-1/2) is.
"""
import math

import numpy as np
import pytest

import oracle_lib as OL


def min_gap(c, r, L=None):
    d = c[:, None, :] - c[None, :, :]
    if L is not None:
        d = d - np.round(d / L) * L
    dist = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(dist, np.inf)
    return (dist - r[:, None] - r[None, :]).min()


def test_generator_is_minimal_standard():
    # 10000 draws of x <- 48271 x mod (2^31 - 1) from x0 = 1 give 399268537 (C++11 minstd_rand's
    # required 10000th value); makeCloud of one sphere in a box draws 3 numbers per attempt
    x = 1
    for _ in range(10000):
        x = (48271 * x) % 2147483647
    assert x == 399268537
    c, r, st, _ = OL.make_cloud((0, 0, 0), (1, 1, 1), rMean=.1, num=1, state=1)
    x = 1
    for _ in range(3):
        x = (48271 * x) % 2147483647
    assert st == x
    u = [(48271 ** k % 2147483647 - 1) / 2147483645. for k in (1, 2, 3)]
    assert np.array_equal(c[0], [.1 + (1 - .2) * u[0], .1 + (1 - .2) * u[1], .1 + (1 - .2) * u[2]])


def test_radius_sequence_rmean():
    num, rMean, fuzz = 200, .004, .25
    c, r, st, _ = OL.make_cloud((0, 0, 0), (.1, .1, .1), rMean=rMean, rRelFuzz=fuzz, num=num, state=11)
    assert len(r) == num
    k = np.arange(num)
    want = rMean * (2 * ((num - k + .5) / (num + 1.) - .5) * fuzz + 1)
    assert np.array_equal(r, want)
    assert np.all(np.diff(r) < 0)
    assert min_gap(c, r) > 0
    assert np.all(c - r[:, None] >= 0) and np.all(c + r[:, None] <= .1)


def test_mean_radius_from_porosity():
    # RDIST_NUM: rMean = (V (1-p) / (4/3 pi (1+f^2) num))^(1/3)
    num, p, f = 100, .8, .1
    c, r, st, _ = OL.make_cloud((0, 0, 0), (.2, .1, .1), rRelFuzz=f, num=num, porosity=p, state=5)
    rm = (.2 * .1 * .1 * (1 - p) / (math.pi * 4 / 3. * (1 + f * f) * num)) ** (1 / 3.)
    k = np.arange(num)
    assert np.allclose(r, rm * (2 * ((num - k + .5) / (num + 1.) - .5) * f + 1), rtol=1e-15, atol=0)
    assert r.mean() == pytest.approx(rm * (1 + f / (num + 1)), rel=1e-14)  # the sequence's mean


def test_periodic_minimum_image_and_psd():
    L = np.array([.05, .06, .07])
    sizes, cumm = [.004, .006, .01], [0., .3, 1.]
    c, r, st, _ = OL.make_cloud((0, 0, 0), L, psdSizes=sizes, psdCumm=cumm, num=150, periodic=True, state=3)
    assert len(r) == 150
    assert np.all(r >= .002 - 1e-15) and np.all(r <= .005 + 1e-15)
    assert min_gap(c, r, L) > 0
    assert np.all(c >= 0) and np.all(c < L + 1e-15)


def test_distribute_mass_pow3():
    # distributeMass: the number density goes as r^-3 on [a, b], so r^-2 is affine in the quantile
    # (pow3Interp, SpherePack.hpp:86): r = a at quantile 0, b at 1
    num, rMean, f = 400, .003, .4
    c, r, st, _ = OL.make_cloud((0, 0, 0), (.1, .1, .1), rMean=rMean, rRelFuzz=f, num=num, distributeMass=True, state=2)
    a, b = rMean * (1 - f), rMean * (1 + f)
    q = (num - np.arange(num) + .5) / (num + 1.)
    assert np.allclose(r ** -2, q * (b ** -2 - a ** -2) + a ** -2, rtol=1e-13, atol=0)
    assert min_gap(c, r) > 0


def test_psd_mass_scaling_fills_the_solid_volume():
    # a PSD linear in diameter between d0 and d1 holds N = 3 V (1-p) / (pi (d1-d0)) (d0^-2 - d1^-2)
    # spheres (the integral of the mass fraction over the sphere volume pi d^3 / 6); asking for more
    # scales the radii by (N/num)^(1/3), and the num spheres then fill V (1-p) of solid
    V, p, d0, d1 = .06 ** 3, .8, .002, .006
    N = 3 * V * (1 - p) / (math.pi * (d1 - d0)) * (d0 ** -2 - d1 ** -2)
    num = int(N * 1.5)
    c, r, st, sc = OL.make_cloud((0, 0, 0), (.06, .06, .06), psdSizes=[d0, d1], psdCumm=[0, 1], num=num, porosity=p,
                                 distributeMass=True, state=4)
    assert len(r) == num
    assert sc == pytest.approx((N / num) ** (1 / 3.), rel=1e-13)
    assert np.all(r >= sc * d0 / 2 * (1 - 1e-15)) and np.all(r <= sc * d1 / 2 * (1 + 1e-15))
    assert (4 / 3. * math.pi * r ** 3).sum() == pytest.approx(V * (1 - p), rel=1e-3)
    assert min_gap(c, r) > 0


def test_retry_with_higher_porosity():
    # 2000 spheres at porosity .3 do not fit a random loose packing: the radii shrink until they do
    c, r, st, _ = OL.make_cloud((0, 0, 0), (.05, .05, .05), num=2000, porosity=.3, state=9)
    assert len(r) == 2000
    rm0 = (.05 ** 3 * .7 / (math.pi * 4 / 3. * 2000)) ** (1 / 3.)
    assert r.mean() < rm0
    assert min_gap(c, r) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["rmean", "num", "psd_periodic", "psd_mass", "stochastic"])
def test_gpu_make_cloud_equals_oracle(case):
    from woodem_amd._native import Context
    ctx = Context(0)
    kw = dict(rmean=dict(mn=(0, 0, 0), mx=(.1, .1, .1), rMean=.003, rRelFuzz=.3, num=3000),
              num=dict(mn=(-.1, 0, .2), mx=(.1, .05, .3), rRelFuzz=.2, num=4000, porosity=.7),
              psd_periodic=dict(mn=(0, 0, 0), mx=(.05, .06, .07), psdSizes=[.002, .004, .006], psdCumm=[0, .4, 1], num=1200,
                                periodic=True),
              psd_mass=dict(mn=(0, 0, 0), mx=(.08, .08, .08), psdSizes=[.002, .003, .008], psdCumm=[0, .5, 1], num=3000,
                            distributeMass=True),
              stochastic=dict(mn=(0, 0, 0), mx=(.07, .07, .07), rMean=.003, rRelFuzz=.4, num=-1))[case]
    mn, mx = kw.pop("mn"), kw.pop("mx")
    cap = 20000
    o_c, o_r, o_st, o_sc = OL.make_cloud(mn, mx, state=12345, cap=cap, **kw)
    st = np.array([12345], np.uint32)
    args = dict(rMean=-1, rRelFuzz=0, num=-1, periodic=False, porosity=.5, psdSizes=[], psdCumm=[], distributeMass=False)
    args.update(kw)
    g_c, g_r, g_sc = ctx.make_cloud(mn, mx, args["rMean"], args["rRelFuzz"], args["num"], args["periodic"], args["porosity"],
                                    args["psdSizes"], args["psdCumm"], args["distributeMass"], st, cap)
    assert len(g_r) == len(o_r) > 100
    assert np.array_equal(g_r, o_r) and np.array_equal(g_c, o_c)
    assert int(st[0]) == o_st and g_sc == o_sc


# ---------------- pack.regularOrtho / regularHexa (py/pack.py:323-350) ----------------
def test_ortho_axes_follow_numpy_arange():
    # the reference lays the grid out with numpy.arange(mn+r, mx-r, 2r+gap) per axis and
    # visits itertools.product(xx, yy, zz): z fastest
    mn, mx, r, gap = np.array([-.013, .2, 1.]), np.array([.1, .31, 1.07]), .004, .0013
    c = OL.regular_pack(0, 0, list(mn) + list(mx), r, gap)
    ax = [np.arange(mn[i] + r, mx[i] - r, 2 * r + gap) for i in range(3)]
    want = np.array([(x, y, z) for x in ax[0] for y in ax[1] for z in ax[2]])
    assert np.array_equal(c, want)  # every lattice point of a box lies inside it with pad r


def test_in_sphere_pad_cancels():
    # inSphere tests |p - c| - pad <= R - pad (py/_packPredicates.cpp:103): the pad cancels, so
    # the spheres are those whose centres lie in the ball (their surfaces may cross it)
    cen, R, r, gap = np.array([.01, -.02, .03]), .05, .004, .001
    c = OL.regular_pack(0, 1, list(cen) + [R], r, gap)
    ax = [np.arange(cen[i] - R + r, cen[i] + R - r, 2 * r + gap) for i in range(3)]
    lat = np.array([(x, y, z) for x in ax[0] for y in ax[1] for z in ax[2]])
    d = np.linalg.norm(lat - cen, axis=1)
    sure = np.abs(d - R) > 1e-12 * R
    inside = set(map(tuple, lat[sure & (d < R)]))
    got = set(map(tuple, c))
    assert inside <= got and not (got & set(map(tuple, lat[sure & (d > R)])))
    assert np.max(np.linalg.norm(c - cen, axis=1)) > R - r  # some do reach out of the ball


def test_in_cylinder_against_axis_projection():
    # inCylinder keeps centres at least r inside both bases and at most R - r from the axis
    c1, c2, R, r, gap = np.array([0., 0., 0.]), np.array([.02, .03, .09]), .025, .003, .0005
    c = OL.regular_pack(0, 2, list(c1) + list(c2) + [R], r, gap)
    assert len(c) > 50
    e = (c2 - c1) / np.linalg.norm(c2 - c1)
    t = (c - c1) @ e
    rad = np.linalg.norm((c - c1) - t[:, None] * e, axis=1)
    ht = np.linalg.norm(c2 - c1)
    tol = 1e-12
    assert np.all(t >= r - tol) and np.all(t <= ht - r + tol) and np.all(rad <= R - r + tol)
    # and every lattice point that satisfies it clearly is kept
    k = np.array([np.sqrt((c1[1] - c2[1]) ** 2 + (c1[2] - c2[2]) ** 2), np.sqrt((c1[0] - c2[0]) ** 2 + (c1[2] - c2[2]) ** 2),
                  np.sqrt((c1[0] - c2[0]) ** 2 + (c1[1] - c2[1]) ** 2)]) / ht
    mn, mx = np.minimum(c1, c2) - R * k, np.maximum(c1, c2) + R * k
    ax = [np.arange(mn[i] + r, mx[i] - r, 2 * r + gap) for i in range(3)]
    lat = np.array([(x, y, z) for x in ax[0] for y in ax[1] for z in ax[2]])
    tl = (lat - c1) @ e
    rl = np.linalg.norm((lat - c1) - tl[:, None] * e, axis=1)
    clear = (tl > r + 1e-9) & (tl < ht - r - 1e-9) & (rl < R - r - 1e-9)
    assert set(map(tuple, lat[clear])) <= set(map(tuple, c))


def test_hexa_neighbour_distances():
    # regularHexa's rows are a sqrt(3)/2 apart and its layers a sqrt(6)/3, alternate rows shifted
    # by a/2 and odd layers by (a/2, hy/2) (py/pack.py:341-348); hy/2 rather than hy/3 makes each
    # inner sphere's 12 nearest neighbours 2 at sqrt(41/48) a, 6 at a and 4 at sqrt(53/48) a
    r, gap = .01, .002
    a = 2 * r + gap
    c = OL.regular_pack(1, 0, [0, 0, 0, .3, .3, .3], r, gap)
    inner = np.all((c > .08) & (c < .22), axis=1)
    d = np.linalg.norm(c[inner][:, None, :] - c[None, :, :], axis=2)
    d.sort(axis=1)
    nn = d[:, 1:13] / a
    want = np.array([np.sqrt(41 / 48.)] * 2 + [1.] * 6 + [np.sqrt(53 / 48.)] * 4)
    assert inner.sum() > 100
    assert np.allclose(nn, want[None, :], rtol=1e-12, atol=0)
    assert np.all(d[:, 13] / a > np.sqrt(53 / 48.) + 1e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("hexa", [0, 1])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_gpu_regular_pack_equals_oracle(hexa, kind):
    from woodem_amd._native import Context
    ctx = Context(0)
    params = {0: [-.05, 0., .1, .07, .09, .2], 1: [.01, -.02, .03, .06], 2: [0., 0., 0., .02, .03, .09, .025]}[kind]
    for r, gap in ((.003, .0005), (.0021, 0.)):
        want = OL.regular_pack(hexa, kind, params, r, gap)
        got = ctx.regular_pack(hexa, kind, params, r, gap)
        assert len(want) > 100
        assert np.array_equal(got, want)
    with pytest.raises(RuntimeError):
        ctx.regular_pack(hexa, kind, params, .003, .0005, cap=10)


@pytest.mark.gpu
def test_gpu_cloud_to_simulation_steps_like_oracle():
    # the Woo scripting path: SpherePack.makeCloud (seeded once per process, as the reference's
    # static generator), toSimulation into a walled box, defaultEngines, Scene.run; the cloud and
    # the trajectories equal the oracle's
    import woodem_amd as woo
    from woodem_amd import pack, utils
    pack.reset_generator()
    sp = pack.SpherePack()
    box = ((0, 0, 0), (.05, .05, .05))
    n = sp.makeCloud(box[0], box[1], rMean=.002, rRelFuzz=.2, num=400, seed=3)
    c, r, st, _ = OL.make_cloud(box[0], box[1], rMean=.002, rRelFuzz=.2, num=400, state=3)
    assert n == 400 and np.array_equal(sp.c, c) and np.array_equal(sp.r, r)
    sp2 = pack.SpherePack()  # the generator goes on where the first cloud left it
    sp2.makeCloud(box[0], box[1], rMean=.002, num=10, seed=99)
    c2, r2, _, _ = OL.make_cloud(box[0], box[1], rMean=.002, num=10, state=st)
    assert np.array_equal(sp2.c, c2)
    S = woo.Scene(fields=[woo.DemField(gravity=(0, 0, -9.81))], dt=2e-5)
    S.dem.par.add(utils.wallBox(box))
    ids = sp.toSimulation(S)
    assert len(ids) == 400
    S.engines = utils.defaultEngines(damping=.2, dynDtPeriod=0)
    S.run(600)
    a = S.dem._compile()
    o = OL.OracleRun(a, S._params(), dict(periodic=False, hSize=np.eye(3), gradV=np.zeros((3, 3)), nextGradV=np.zeros((3, 3)),
                                          homoDeform=4))
    o.step(600)
    no = o.nodes()
    for i in ids[::7]:
        assert np.array_equal(S.dem.par[i].pos, no["pos"][i])
    assert S.dem.con.countReal() == int(o.contacts()["isReal"].sum()) > 0
    pack.reset_generator()


# ---------------- SpherePack text files (SpherePack.cpp:73-129) ----------------
def test_sphere_pack_file_format(tmp_path):
    from woodem_amd import pack
    sp = pack.SpherePack()
    sp.c = np.array([[0.0123456789, 1e-5, 2.], [1., 2.5, 3.25]])
    sp.r = np.array([0.001, 1234567.])
    sp.cellSize = np.array([1., 2., 3.5])
    sp.userData = "made by a test"
    f = str(tmp_path / "a.txt")
    sp.save(f)
    # C++ streams' default format: 6 significant digits, exponent form where %g uses it
    assert open(f).read() == ("##PERIODIC:: 1 2 3.5\n##USER-DATA:: made by a test\n"
                              "0.0123457 1e-05 2 0.001\n1 2.5 3.25 1.23457e+06\n")
    sp2 = pack.SpherePack()
    sp2.load(f)
    assert sp2.userData == "made by a test" and np.array_equal(sp2.cellSize, [1., 2., 3.5])
    assert np.array_equal(sp2.c, [[0.0123457, 1e-5, 2.], [1., 2.5, 3.25]]) and np.array_equal(sp2.r, [0.001, 1.23457e6])
    assert not sp2.hasClumps()
    g = tmp_path / "b.txt"
    g.write_text("0 0 0 1 3\n\t1 1 1 .5 3\n\n2 2 2 .25 -1\n")  # tabs and blank lines are skipped
    sp3 = pack.SpherePack()
    sp3.loadTxt(str(g))
    assert list(sp3.clumpId) == [3, 3, -1] and sp3.hasClumps()
    sp3.saveTxt(str(tmp_path / "c.txt"))
    assert open(str(tmp_path / "c.txt")).read() == "0 0 0 1 3\n1 1 1 0.5 3\n2 2 2 0.25 -1\n"
    bad = tmp_path / "d.txt"
    bad.write_text("0 0 0 1\n1 1 1 1 2\n")
    with pytest.raises(ValueError):
        sp3.load(str(bad))
    bad.write_text("##PERIODIC:: 1 2\n")
    with pytest.raises(ValueError):
        sp3.load(str(bad))
    with pytest.raises(ValueError):
        sp3.load(str(tmp_path / "missing.txt"))


def test_pack_refusals_before_any_device_call():
    # the host checks that run before the library is asked for anything
    from woodem_amd import pack
    with pytest.raises(NotImplementedError):
        pack.regularOrtho(object(), .01, 0.)  # a predicate outside inAlignedBox/inSphere/inCylinder
    sp = pack.SpherePack()
    with pytest.raises(NotImplementedError):
        sp.makeCloud((0, 0, 0), (1, 1, 1), .1, hSize=np.diag([1., 1., 1.]) + .1)  # sheared cells
    sp.c, sp.r = np.zeros((1, 3)), np.ones(1)
    with pytest.raises(NotImplementedError):
        sp.makeCloud((0, 0, 0), (1, 1, 1), .1)  # into a non-empty pack
    sp.clumpId = np.array([0])
    with pytest.raises(NotImplementedError):
        sp.toSimulation(None)  # clumps
    b = pack.inAlignedBox((0, 0, 0), (1, 2, 3))
    assert b.kind == 0 and b.params == (0., 0., 0., 1., 2., 3.) and b.aabb() == ((0., 0., 0.), (1., 2., 3.))
    assert pack.inSphere((1, 2, 3), .5).params == (1., 2., 3., .5)
    assert pack.inCylinder((0, 0, 0), (0, 0, 1), .2).params == (0., 0., 0., 0., 0., 1., .2)


# ---------------- predicates on the host, SpherePack.filtered, randomLoosePsd ----------------
def _ortho_lattice(mn, mx, r, gap):
    ax = [np.arange(mn[i] + r, mx[i] - r, 2 * r + gap) for i in range(3)]
    return np.array([(x, y, z) for x in ax[0] for y in ax[1] for z in ax[2]])


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_host_predicates_select_what_the_oracle_lattice_keeps(kind):
    # Predicate.contains with pad = r picks out of the full lattice exactly the points the
    # oracle's regularOrtho keeps (whose predicates the geometric tests above pin), in order
    from woodem_amd import pack
    r, gap = .0031, .0007
    pred = [pack.inAlignedBox((-.02, .01, .1), (.03, .05, .13)), pack.inSphere((.01, -.02, .03), .041),
            pack.inCylinder((0., .01, -.01), (.02, .03, .07), .022)][kind]
    mn, mx = (np.asarray(x) for x in pred.aabb())
    lat = _ortho_lattice(mn, mx, r, gap)
    keep = pred.contains(lat, r)
    assert keep.sum() > 50 and (kind == 0 or keep.sum() < len(lat))  # a box keeps its whole lattice
    assert np.array_equal(lat[keep], OL.regular_pack(0, kind, list(pred.params), r, gap))
    assert pred(lat[int(np.argmax(keep))], r)
    if kind:
        assert not pred(lat[int(np.argmin(keep))], r)
    else:
        assert not pred(lat[0] - 1e-9, r) and pred(mx - r, r) and not pred(mx - r + 1e-9, r)


def test_predicate_aabbs():
    from woodem_amd import pack
    mn, mx = pack.inSphere((1., 2., 3.), .5).aabb()
    assert mn == (.5, 1.5, 2.5) and mx == (1.5, 2.5, 3.5)
    # a cylinder's aabb is that of its two base discs: a disc of radius R with unit normal e
    # reaches R sqrt(1 - e_d^2) along axis d; compare with points on the rims
    A, B, R = np.array([.1, -.2, .3]), np.array([.4, .2, -.1]), .07
    mn, mx = (np.asarray(x) for x in pack.inCylinder(A, B, R).aabb())
    e = (B - A) / np.linalg.norm(B - A)
    u = np.cross(e, [1., 0., 0.])
    u /= np.linalg.norm(u)
    v = np.cross(e, u)
    t = np.linspace(0, 2 * np.pi, 20001)
    rim = (R * (np.cos(t)[:, None] * u + np.sin(t)[:, None] * v))
    pts = np.concatenate([A + rim, B + rim])
    assert np.allclose(mn, pts.min(0), atol=1e-8) and np.allclose(mx, pts.max(0), atol=1e-8)
    assert np.all(mn <= pts.min(0) + 1e-15) and np.all(mx >= pts.max(0) - 1e-15)
    with pytest.raises(ValueError):
        pack.inCylinder((1, 1, 1), (1, 1, 1), .1)


def _pack_of(c, r, cid=None):
    from woodem_amd import pack
    sp = pack.SpherePack()
    sp.c, sp.r = np.asarray(c, np.float64), np.asarray(r, np.float64)
    sp.clumpId = np.full(len(sp.r), -1, np.int64) if cid is None else np.asarray(cid, np.int64)
    return sp


def test_filtered_recenters_and_pads_by_radius(caplog):
    from woodem_amd import pack
    rng = np.random.default_rng(3)
    c = rng.uniform(0, 1, (400, 3))
    r = rng.uniform(.01, .05, 400)
    sp = _pack_of(c, r)
    sp.cellSize = np.array([1., 1., 1.])
    box = pack.inAlignedBox((10.2, 10.3, 10.1), (10.8, 10.7, 10.9))
    f = sp.filtered(box)
    smn, smx = sp.aabb()
    off = np.array([10.5, 10.5, 10.5]) - (smn + smx) / 2  # the packing's aabb centre goes to the box's
    cc = c + off
    inside = np.all((cc - r[:, None] >= box.params[:3]) & (cc + r[:, None] <= box.params[3:]), axis=1)  # fully inside
    assert np.array_equal(f.c, cc[inside]) and np.array_equal(f.r, r[inside]) and 20 < len(f) < 400
    assert np.array_equal(f.cellSize, sp.cellSize) and not f.hasClumps()
    # without recentering nothing lies in a box that far away
    assert len(sp.filtered(box, recenter=False)) == 0
    # a predicate larger than the packing is warned about, the spheres are all kept
    with caplog.at_level("WARNING"):
        g = sp.filtered(pack.inAlignedBox((-5, -5, -5), (5, 5, 5)), recenter=False)
    assert len(g) == 400 and "outside of the predicate aabb" in caplog.text


def test_filtered_removes_whole_clumps():
    from woodem_amd import pack
    # clump 0 has one sphere outside the sphere predicate, clump 1 is inside, two loose spheres
    c = [(0, 0, 0), (.9, 0, 0), (1.2, 0, 0), (0, .2, 0), (0, .3, 0), (0, -.4, 0), (0, 0, 1.5)]
    sp = _pack_of(c, [.1] * 7, [0, 0, 0, 1, 1, -1, -1])
    f = sp.filtered(pack.inSphere((0, 0, 0), 1.), recenter=False)
    assert np.array_equal(f.c, np.array(c, float)[[3, 4, 5]]) and list(f.clumpId) == [1, 1, -1]
    assert len(pack.filterSpherePack(pack.inSphere((0, 0, 0), 1.), sp)) == len(sp.filtered(pack.inSphere((0, 0, 0), 1.)))


def test_sanitize_psd():
    from woodem_amd import pack
    # leading zero points but the last, trailing top points but the first, passing scaled to 1
    got = pack.sanitizePsd([(.1, 0), (.2, 0), (.3, 0), (.4, 20), (.5, 50), (.6, 50), (.7, 50)])
    assert got == [(.3, 0.), (.4, .4), (.5, 1.), ]
    assert pack.sanitizePsd([(.1, .2), (.2, 1.)]) == [(.1, .2), (.2, 1.)]
    for bad in ([(.2, 0), (.1, 1)], [(.1, .5), (.2, .4)], [(-.1, 0), (.2, 1)]):
        with pytest.raises(RuntimeError):
            pack.sanitizePsd(bad)


def test_random_loose_psd_refusals():
    from woodem_amd import pack
    box = pack.inAlignedBox((0, 0, 0), (1, 1, 1))
    with pytest.raises(ValueError):
        pack.randomLoosePsd(box, [(.1, 0), (.2, 1)])  # returnSpherePack must be True, as in the reference
    with pytest.raises(NotImplementedError):
        pack.randomLoosePsd(box, [(.1, 0), (.2, 1)], clumps=[object()], returnSpherePack=True)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["rmean_mass", "psd_mass"])
def test_gpu_stochastic_mass_cloud_within_an_ulp_of_the_oracle(case):
    # with num < 0 the device draws each radius as it goes; by mass that is pow(t, -1/2), which
    # the device rounds correctly and the C library's pow (the oracle's, as the reference's)
    # misrounds by an ulp now and then: every sphere is placed alike, the radii agree to an
    # ulp, and the centres to the few ulps that such a radius moves them
    from woodem_amd._native import Context
    ctx = Context(0)
    kw = dict(rmean_mass=dict(mn=(0, 0, 0), mx=(.06, .06, .06), rMean=.003, rRelFuzz=.4, distributeMass=True),
              psd_mass=dict(mn=(0, 0, 0), mx=(.06, .05, .07), psdSizes=[.002, .003, .006], psdCumm=[0, .3, 1], distributeMass=True))[case]
    mn, mx = kw.pop("mn"), kw.pop("mx")
    o_c, o_r, o_st, _ = OL.make_cloud(mn, mx, num=-1, state=4242, cap=60000, **kw)
    st = np.array([4242], np.uint32)
    args = dict(rMean=-1, rRelFuzz=0, psdSizes=[], psdCumm=[])
    args.update(kw)
    g_c, g_r, _ = ctx.make_cloud(mn, mx, args["rMean"], args["rRelFuzz"], -1, False, .5, args["psdSizes"], args["psdCumm"], True, st, 60000)
    assert len(g_r) == len(o_r) > 500 and int(st[0]) == o_st
    dr = np.abs(g_r - o_r)
    assert np.all(dr <= np.spacing(o_r)) and np.count_nonzero(dr) <= .01 * len(o_r)
    assert np.max(np.abs(g_c - o_c)) <= 1e-16


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_gpu_random_loose_psd_is_the_oracle_cloud_filtered(kind):
    # randomLoosePsd is the device cloud of the PSD by mass in the predicate's aabb, filtered;
    # that cloud is the oracle's to an ulp (previous test).  Filtered, every sphere lies in the
    # predicate, none overlap, the radii stay in the PSD's range and the mass below the middle
    # diameter is near the PSD's passing there.
    from woodem_amd import pack
    from woodem_amd._native import Context
    ctx = Context(0)
    pred = [pack.inAlignedBox((0, 0, 0), (.06, .05, .04)), pack.inSphere((.1, .2, .3), .035),
            pack.inCylinder((0, 0, 0), (.01, .02, .06), .025)][kind]
    psd = [(.002, 0.), (.003, 30.), (.005, 100.)]  # percent passing, as a sieve curve is often given
    pack.reset_generator()
    sp = pack.randomLoosePsd(pred, psd, returnSpherePack=True, seed=777, ctx=ctx)
    mn, mx = pred.aabb()
    pack.reset_generator()
    cl = pack.SpherePack()
    cl.makeCloud(mn, mx, psdSizes=[.002, .003, .005], psdCumm=[0., .3, 1.], distributeMass=True, seed=777, ctx=ctx)
    f = cl.filtered(pred)
    assert np.array_equal(sp.c, f.c) and np.array_equal(sp.r, f.r)
    o_c, o_r, _, _ = OL.make_cloud(mn, mx, psdSizes=[.002, .003, .005], psdCumm=[0., .3, 1.], distributeMass=True, state=777,
                                   cap=len(cl) + 20000)
    assert len(o_r) == len(cl) and np.all(np.abs(cl.r - o_r) <= np.spacing(o_r))
    want = _pack_of(o_c, o_r).filtered(pred)
    assert len(sp) == len(want) > 200
    assert np.max(np.abs(sp.c - want.c)) <= 1e-15 and np.all(np.abs(sp.r - want.r) <= np.spacing(want.r))
    assert np.all(pred.contains(sp.c, sp.r))
    assert min_gap(sp.c, sp.r) >= -1e-12
    assert sp.r.min() >= .001 - 1e-15 and sp.r.max() <= .0025 + 1e-15
    m = sp.r ** 3
    assert abs(m[sp.r <= .0015].sum() / m.sum() - .3) < .1
