"""Woo-style host API (woodem_amd.core/dem/utils): scene compilation into the ABI arrays and
engine attributes into wdm_params, following py/utils.py:38-268 and the engine attribute lists.
CPU-only except the last test."""
import math

import numpy as np
import pytest

import woodem_amd as woo
from woodem_amd import soa, utils
from woodem_amd._native import LAW_HERTZ_DMT, LAW_IDEALELPL, SHAPE_SPHERE, SHAPE_WALL


def small_scene():
    S = woo.Scene(fields=[woo.DemField(gravity=(0, 0, -9.81))], dt=2.6e-5)
    S.dem.par.add(utils.wallBox(((0, 0, 0), (.3, .25, .25))))
    S.dem.par.add([utils.sphere((.1, .1, z), .005) for z in (.0055, .02, .04)])
    S.engines = utils.defaultEngines(damping=.2, dynDtPeriod=0)
    return S


def test_wallbox_senses_and_masks():
    # py/utils.py:201-210: sense +1 on minimum faces, -1 on maximum faces; masks py/utils.py:107,177
    S = small_scene()
    a = S.dem._compile()
    walls = a.shape == SHAPE_WALL
    assert walls.sum() == 6
    assert list(a.axis[walls]) == [0, 1, 2, 0, 1, 2]
    assert list(a.sense[walls]) == [1, 1, 1, -1, -1, -1]
    assert np.all(a.mask[walls] == soa.DEFAULT_BOUNDARY_MASK)
    assert np.all(a.mask[a.shape == SHAPE_SPHERE] == soa.DEFAULT_MOVABLE_MASK)
    assert np.all(a.flags[walls] & soa.DOF_ALL == soa.DOF_ALL)
    assert np.all(a.nodes[:, 0] == np.arange(a.n_par))


def test_sphere_mass_from_default_material():
    # defaultMaterial: FrictMat(density=1e3, young=1e7, ktDivKn=.2, tanPhi=tan(.5)), py/utils.py:48-53
    p = utils.sphere((0, 0, 0), .01)
    assert p.material.density == 1e3 and p.material.tanPhi == math.tan(.5)
    assert p.nodes[0].dem.mass == pytest.approx(4 / 3. * math.pi * 1e-6 * 1e3, rel=1e-15)


def test_params_from_default_engines():
    S = small_scene()
    p = S._params()
    assert p.law == LAW_IDEALELPL and p.damping == .2 and p.verletDist == -.05
    assert tuple(p.gravity) == (0, 0, -9.81)
    assert p.loneMask == soa.DEFAULT_LONE_MASK and p.trsfRenorm == 100 and p.dist00 == 1


def test_hertz_engines_params():
    S = small_scene()
    S.engines = utils.defaultEngines(damping=0., dynDtPeriod=0, cp2=woo.Cp2_FrictMat_HertzPhys(poisson=.3, en=.7),
                                     law=woo.Law2_L6Geom_HertzPhys_DMT())
    p = S._params()
    assert p.law == LAW_HERTZ_DMT and p.poisson == .3 and p.en == .7


def test_unsupported_configurations_raise():
    S = small_scene()
    S.engines = utils.defaultEngines(damping=.2)  # DynDt included by default (dynDtPeriod=100)
    S._params()
    S.engines[-1].realPeriod = 5.  # wall-clock periodicity cannot be known ahead of a window
    with pytest.raises(NotImplementedError):
        S.engines[-1]._native(.9)
    S.engines = utils.defaultEngines(damping=.2)
    S.engines = [S.engines[-1]] + S.engines[:-1]  # DynDt must follow ContactLoop
    with pytest.raises(ValueError):
        S._params()
    S.engines = utils.defaultEngines(damping=.2, dynDtPeriod=0)
    S.engines[2].lawFunctors = [woo.Law2_L6Geom_HertzPhys_DMT()]
    with pytest.raises(ValueError):
        S._params()
    S.engines = utils.defaultEngines(damping=.2, dynDtPeriod=0)
    S.engines[0].reset = False
    with pytest.raises(ValueError):
        S._params()


def test_auto_dt_is_pwave():
    S = small_scene()
    S.dt = float("nan")
    p = S._params()
    assert p.dt == pytest.approx(.9 * utils.spherePWaveDt(.005, 1e3, 1e7), rel=1e-15)


@pytest.mark.gpu
def test_scene_runs_on_gpu_and_matches_oracle():
    import oracle_lib as OL
    from woodem_amd import runner
    S = small_scene()
    S.trackEnergy = True
    S.run(1500)
    a = S.dem._compile()  # initial state (particles are unchanged descriptions)
    o = OL.OracleRun(a, S._params(), dict(periodic=False, hSize=np.eye(3), gradV=np.zeros((3, 3)),
                                          nextGradV=np.zeros((3, 3)), homoDeform=4))
    o.step(1500)
    no = o.nodes()
    for i in range(a.n_par):
        assert np.array_equal(S.dem.par[i].pos, no["pos"][i])
    e = S.energy
    vo, co = o.energy()
    for k, name in enumerate(OL.ENERGY_NAMES):
        assert (name in e) == bool(co[k])
        if name in e:
            assert e[name] == pytest.approx(vo[k], rel=1e-9, abs=1e-15)
    assert S.dem.con.countReal() == int(o.contacts()["isReal"].sum())
