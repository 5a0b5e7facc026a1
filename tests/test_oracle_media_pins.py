"""Pins for the oracle's heterogeneous medium and volumetric path integrator:
closed forms that a dropped term, a wrong cell
size, an unskipped empty cell or a biased estimator would fail.

* Grid DDA with several voxels per macro cell: alternate slabs of macro cells
  hold a constant density, the others are empty.  Ratio tracking
  (heterogeneous.cpp:106-229) must average exp(-tau), and delta tracking
  (heterogeneous.cpp:251-376) must leave the segment without a collision
  with probability exp(-tau) and place collisions with the CDF
  1 - exp(-tau(t)).  tau is sigma_t times the integral of the trilinear
  field over the cells whose majorant is positive: the reference builds a
  cell's majorant from the voxels whose integer coordinate lies in it
  (heterogeneous.cpp:84-101) and its DDA skips cells with a zero majorant, so
  the rising ramp of the trilinear field in the last voxel-width of an empty
  cell is not seen.  Each test also checks that the same statistic rejects the
  tau that would count that ramp.
* Volumetric furnace (volPath.cpp:15-298): a camera inside a closed box whose
  Lambertian walls (albedo a) emit Le, with a non-absorbing medium behind an
  index boundary (a homogeneous sphere, or a density grid in a box): every
  pixel converges to Le / (1 - a), since a medium of albedo 1 changes nothing
  in an isotropic equilibrium.  Medium sampling, phase sampling, NEE through
  the medium with transmittance, the MIS weights, surface and volume Russian
  roulette and the null-BSDF crossings all enter."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import stats as sst

from gianduia_amd import _abi as A
import oracle as O
from oracle import RefScene, fp, vec
from scenes import SceneBuilder, trs

N_GRID = 64


def _slab_density(macro_res, rho):
    """64^3 grid: macro cells along x alternate between density rho (even
    cells) and 0 (odd cells); x is the fastest axis in memory."""
    cell = N_GRID // macro_res
    dens = np.zeros((N_GRID, N_GRID, N_GRID), np.float32)
    for k in range(0, macro_res, 2):
        dens[:, :, k * cell:(k + 1) * cell] = rho
    return dens


def _grid_scene(dens, macro_res, sigma_t):
    """A scene holding one grey, non-absorbing grid medium whose index space
    is world space (to_local = identity)."""
    b = SceneBuilder()
    mt = b.matte()
    b.prim(-1, material=mt, sphere_radius=0.01, radiance=(1, 1, 1))
    d = b.desc()
    med = (A.gnd_medium_desc * 1)()
    med[0].type = A.GND_MEDIUM_GRID
    med[0].sigma_a[:] = (0.0, 0.0, 0.0)
    med[0].sigma_s[:] = (sigma_t, sigma_t, sigma_t)
    med[0].nx = med[0].ny = med[0].nz = dens.shape[0]
    dens = np.ascontiguousarray(dens)
    b.keep += [dens, med]
    med[0].density = dens.ctypes.data_as(C.POINTER(C.c_float))
    med[0].density_scale = 1.0
    med[0].macro_res = macro_res
    eye = np.eye(4, dtype=np.float32).reshape(-1).tolist()
    med[0].to_local[:] = eye
    med[0].to_local_inv[:] = eye
    d.nmedia = 1
    d.media = med
    return RefScene(d)


def _voxels_x(dens, y, z):
    """Voxel values along x at the integer (y, z) of a slab grid (constant in y, z)."""
    return dens[int(z), int(y), :].astype(np.float64)


def _tau(vx, macro_res, o, d, t0, t1, count_empty_ramps=False):
    """sigma_t-free optical depth along o + t d, t in [t0, t1], of the
    trilinear field of voxel values vx along x (voxels outside are 0), over
    the macro cells with a positive majorant (or, with count_empty_ramps,
    over every cell).  The field is linear between integers, so splitting at
    integers and integrating each piece exactly is exact."""
    n = len(vx)
    cell = n // macro_res
    maj = np.array([vx[c * cell:(c + 1) * cell].max() > 0 for c in range(macro_res)])

    def rho(x):
        i = math.floor(x)
        u = x - i
        a = vx[i] if 0 <= i < n else 0.0
        b = vx[i + 1] if 0 <= i + 1 < n else 0.0
        return a + u * (b - a)

    xa, xb = o[0] + d[0] * t0, o[0] + d[0] * t1
    lo, hi = min(xa, xb), max(xa, xb)
    cuts = [lo] + [float(k) for k in range(math.floor(lo) + 1, math.ceil(hi))] + [hi]
    acc = 0.0
    for a, b in zip(cuts[:-1], cuts[1:]):
        if b <= a:
            continue
        c = min(max(int(math.floor(0.5 * (a + b) / cell)), 0), macro_res - 1)
        if not (count_empty_ramps or maj[c]):
            continue
        acc += 0.5 * (rho(a) + rho(b)) * (b - a)
    return acc / abs(d[0])


RAYS = [  # origin, unnormalised direction, tMax: along x, and diagonal in x, y and z
    ((0.25, 20.5, 30.5), (1.0, 0.0, 0.0), 59.5),
    ((2.0, 10.25, 50.75), (1.0, 0.37, -0.29), 55.0),
    ((62.5, 40.5, 12.5), (-1.0, -0.21, 0.33), 58.0),
]


@pytest.mark.parametrize("macro_res", [8, 16])
@pytest.mark.parametrize("ray", range(len(RAYS)))
def test_grid_ratio_tracking_multi_voxel_cells(macro_res, ray):
    rho, sig = 0.5, 0.1
    dens = _slab_density(macro_res, rho)
    s = _grid_scene(dens, macro_res, sig)
    o, d, tmax = RAYS[ray]
    vx = _voxels_x(dens, o[1], o[2])
    tau = sig * _tau(vx, macro_res, o, d, 1e-4, tmax)
    tau_ramps = sig * _tau(vx, macro_res, o, d, 1e-4, tmax, count_empty_ramps=True)
    out = np.zeros(7, np.float32)
    n = 20000
    vals = np.empty(n)
    for seed in range(n):
        O.lib().gnd_ref_medium(s.h, 0, 0, fp(vec(*o)), fp(vec(*d)), tmax, seed, fp(out))
        vals[seed] = out[0]
    mean, se = vals.mean(), vals.std() / math.sqrt(n)
    assert abs(mean - math.exp(-tau)) < 4 * se + 1e-3, (mean, math.exp(-tau), se)
    # the same statistic rejects the optical depth that counts the skipped ramps
    assert abs(mean - math.exp(-tau_ramps)) > 4 * se + 1e-3


@pytest.mark.parametrize("macro_res", [8, 16])
@pytest.mark.parametrize("ray", range(len(RAYS)))
def test_grid_delta_tracking_multi_voxel_cells(macro_res, ray):
    rho, sig = 0.5, 0.1
    dens = _slab_density(macro_res, rho)
    s = _grid_scene(dens, macro_res, sig)
    o, d, tmax = RAYS[ray]
    vx = _voxels_x(dens, o[1], o[2])
    tau = sig * _tau(vx, macro_res, o, d, 1e-4, tmax)
    out = np.zeros(7, np.float32)
    n = 20000
    ts = []
    for seed in range(n):
        O.lib().gnd_ref_medium(s.h, 0, 1, fp(vec(*o)), fp(vec(*d)), tmax, seed, fp(out))
        if out[3] > 0:
            ts.append((out[4] - o[0]) / d[0])
            # a collision returns 1 / sigma_t(p) (grey: the same in every channel)
            assert out[0] > 0
    p_free = 1.0 - len(ts) / n
    se = math.sqrt(p_free * (1 - p_free) / n)
    assert abs(p_free - math.exp(-tau)) < 4 * se, (p_free, math.exp(-tau))
    # collision distances follow F(t) = (1 - exp(-tau(t))) / (1 - exp(-tau(tMax)))
    norm = 1.0 - math.exp(-tau)
    u = np.array([(1.0 - math.exp(-sig * _tau(vx, macro_res, o, d, 1e-4, t))) / norm for t in ts])
    assert sst.kstest(u, "uniform").pvalue > 1e-3
    ramps = lambda t: sig * _tau(vx, macro_res, o, d, 1e-4, t, count_empty_ramps=True)  # noqa: E731
    norm_r = 1.0 - math.exp(-ramps(tmax))
    u_r = np.array([(1.0 - math.exp(-ramps(t))) / norm_r for t in ts])
    assert sst.kstest(u_r, "uniform").pvalue < 1e-3 or abs(p_free - math.exp(-ramps(tmax))) > 4 * se


# ---------------------------------------------------------------- volumetric furnace
def _outward_box(center, half):
    """The 12 triangles of an axis-aligned box with outward normals."""
    tris = []
    c, h = np.asarray(center, np.float32), np.asarray(half, np.float32)
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            u, v = [a for a in range(3) if a != axis]
            corners = []
            for a, b in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                p = [0.0, 0.0, 0.0]
                p[axis], p[u], p[v] = sgn, a, b
                corners.append(c + h * np.array(p, np.float32))
            q = np.array(corners, np.float32)
            out = np.zeros(3, np.float32)
            out[axis] = sgn
            for t in ((0, 1, 2), (0, 2, 3)):
                tri = q[list(t)]
                if np.dot(np.cross(tri[1] - tri[0], tri[2] - tri[0]), out) < 0:
                    tri = tri[[0, 2, 1]]
                tris.append(tri.reshape(9))
    return np.array(tris, np.float32)


def _furnace_scene(kind, albedo):
    from test_oracle_pins import _inward_box
    tris, nrm = _inward_box()
    b = SceneBuilder()
    wall = b.matte((albedo, albedo, albedo))
    b.prim(b.mesh(tris, nrm), material=wall, radiance=(1.0, 1.0, 1.0))
    med = (A.gnd_medium_desc * 1)()
    eye = np.eye(4, dtype=np.float32)
    if kind == "homogeneous":
        med[0].type = A.GND_MEDIUM_HOMOGENEOUS
        med[0].sigma_a[:] = (0.0, 0.0, 0.0)
        med[0].sigma_s[:] = (3.0, 3.0, 3.0)
        med[0].to_local[:] = eye.reshape(-1).tolist()
        M, Mi = trs(trans=(0.0, 0.0, -0.55))
        b.prim(-1, M, Mi, b.index(0), sphere_radius=0.4)
    else:
        n = 8
        dens = np.ones((n, n, n), np.float32)
        b.keep.append(dens)
        med[0].type = A.GND_MEDIUM_GRID
        med[0].sigma_a[:] = (0.0, 0.0, 0.0)
        med[0].sigma_s[:] = (4.0, 4.0, 4.0)
        med[0].nx = med[0].ny = med[0].nz = n
        med[0].density = dens.ctypes.data_as(C.POINTER(C.c_float))
        med[0].density_scale = 1.0
        med[0].macro_res = 4
        # box [-0.5,0.5] x [-0.5,0.5] x [-1.2,-0.2] -> index [0,8]^3
        lo = np.array([-0.5, -0.5, -1.2], np.float32)
        T = np.diag([n / 1.0, n / 1.0, n / 1.0, 1.0]).astype(np.float32)
        T[:3, 3] = -lo * n
        med[0].to_local[:] = T.T.reshape(-1).tolist()
        med[0].to_local_inv[:] = np.linalg.inv(T).T.reshape(-1).tolist()
        b.prim(b.mesh(_outward_box((0.0, 0.0, -0.7), (0.5, 0.5, 0.5))), material=b.index(0))
    b.keep.append(med)
    d = b.desc(width=8, height=8)
    d.nmedia = 1
    d.media = med
    d.integrator = A.GND_INTEGRATOR_VOLPATH
    d.max_depth = 1000  # volPath.cpp:12 default
    return d


@pytest.mark.parametrize("kind", ["homogeneous", "grid"])
@pytest.mark.parametrize("albedo", [0.5, 0.8])
def test_volpath_furnace(kind, albedo):
    d = _furnace_scene(kind, albedo)
    ref = RefScene(d)
    spp = 512
    rgbw, _ = ref.render(threads=8, spp=spp, width=8, height=8)
    img = rgbw[..., :3] / rgbw[..., 3:4]
    expect = 1.0 / (1.0 - albedo)
    assert abs(img.mean() - expect) < 0.02 * expect, img.mean()
    assert np.all(np.abs(img - expect) < 0.25 * expect)
