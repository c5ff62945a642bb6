"""Oracle geometry and physics helpers (CPU), each pinned by brute force or a closed form.

- photontrack_get_grid_coord (src/mc3d_gpu.cu:53-63): the voxel index is i for i*dx < x <
  (i+1)*dx (float products), and on a face x == i*dx it is i when moving up (ux >= 0), i-1 when
  moving down - checked against brute force over faces, their float neighbours and interiors.
- photontrack_intersection (:29-49): distance to the nearest face (float64 reference), exact snap
  of the crossed coordinate to the face.
- Fresnel (:227-245): in the exact mode the textbook unpolarised reflectance from Snell's law; in
  the reference mode (R2 denominator cos_ai*cos_at + sin_at^2, :241) the same normal-incidence
  limit ((n1-n2)/(n1+n2))^2, a different value at oblique incidence, NaN at exactly normal
  incidence (so the packet transmits) and total internal reflection beyond the critical angle.
- Henyey-Greenstein (:282-283): the sampled cos(theta) inverts the HG cumulative distribution and
  its mean is g.
- photontrack_calc_specular_reflection (:67-70): ((n0-n1)/(n0+n1))^2.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O


def lib():
    return O.lib()


@pytest.mark.parametrize("dx", [0.01, 0.002, 0.1, 0.0137])
def test_grid_coord_brute_force(dx):
    dx = np.float32(dx)
    for i in range(-3, 70):
        face = np.float32(np.float32(i) * dx)
        nxt = np.float32(np.float32(i + 1) * dx)
        pts = [face, np.nextafter(face, np.float32(np.inf)), np.nextafter(face, np.float32(-np.inf)),
               np.float32(face + np.float32(0.37) * (nxt - face))]
        for x in pts:
            for ux in (1.0, -1.0, 0.0, -0.0, 0.3):
                got = lib().ymc_oracle_grid_coord(float(x), ux, float(dx))
                # brute force: the voxel v with v*dx <= x < (v+1)*dx in float products
                v = next(v for v in range(i - 3, i + 4)
                         if np.float32(np.float32(v) * dx) <= x < np.float32(np.float32(v + 1) * dx))
                if x == np.float32(np.float32(v) * dx) and not (ux >= 0):
                    v -= 1
                assert got == v, (float(x), ux, float(dx), got, v)


def test_intersection_distance_and_snap():
    rng = np.random.default_rng(5)
    d = np.array([0.01, 0.02, 0.005], np.float32)
    for _ in range(3000):
        ijk = rng.integers(0, 20, 3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pos = ((ijk + rng.uniform(0.01, 0.99, 3)) * d).astype(np.float32)
        uf = u.astype(np.float32)
        p = pos.copy()
        t = lib().ymc_oracle_intersection(int(ijk[0]), int(ijk[1]), int(ijk[2]), *map(float, d),
                                          p.ctypes.data_as(C.POINTER(C.c_float)),
                                          uf.ctypes.data_as(C.POINTER(C.c_float)))
        faces = (np.where(uf >= 0, ijk + 1, ijk).astype(np.float32) * d).astype(np.float32)  # float(i+1)*dx
        tt = (faces.astype(np.float64) - pos.astype(np.float64)) / uf.astype(np.float64)
        assert t == pytest.approx(tt.min(), rel=1e-6, abs=1e-12)
        ax = int(np.argmin(tt))
        assert p[ax] == faces[ax]  # snapped to the face exactly
        for o in range(3):
            if o != ax:
                assert p[o] == pytest.approx(pos[o] + uf[o] * t, abs=1e-6)


def fresnel(cos_ai, ni, nt, mode):
    cat = C.c_float()
    return lib().ymc_oracle_fresnel(cos_ai, ni, nt, mode, C.byref(cat)), cat.value


def textbook(cos_i, n1, n2):
    sin_t = n1 / n2 * np.sqrt(1 - cos_i ** 2)
    if sin_t >= 1:
        return 1.0
    cos_t = np.sqrt(1 - sin_t ** 2)
    rs = ((n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)) ** 2
    rp = ((n1 * cos_t - n2 * cos_i) / (n1 * cos_t + n2 * cos_i)) ** 2
    return 0.5 * (rs + rp)


@pytest.mark.parametrize("ni,nt", [(1.4, 1.0), (1.0, 1.4), (1.44, 1.39), (1.33, 1.5)])
def test_fresnel_exact_mode_is_textbook(ni, nt):
    for cos_i in np.linspace(0.02, 0.999, 60):
        R, _ = fresnel(cos_i, ni, nt, 1)
        assert R == pytest.approx(textbook(cos_i, ni, nt), abs=2e-5), cos_i


@pytest.mark.parametrize("ni,nt", [(1.4, 1.0), (1.0, 1.4), (1.44, 1.39)])
def test_fresnel_reference_mode(ni, nt):
    normal = ((ni - nt) / (ni + nt)) ** 2
    R, _ = fresnel(0.99999, ni, nt, 0)
    assert R == pytest.approx(normal, rel=2e-3)
    R, cat = fresnel(1.0, ni, nt, 0)
    assert np.isnan(R) and cat == 1.0  # 0/0 at exact normal incidence: the packet transmits
    diffs = [abs(fresnel(c, ni, nt, 0)[0] - fresnel(c, ni, nt, 1)[0]) for c in np.linspace(0.3, 0.9, 13)]
    assert max(diffs) > 1e-4  # the R2 denominator of src/mc3d_gpu.cu:241 changes oblique values
    if ni > nt:
        crit = np.sqrt(1 - (nt / ni) ** 2)
        assert fresnel(crit * 0.98, ni, nt, 0)[0] == 1.0
        assert fresnel(crit * 0.98, ni, nt, 1)[0] == 1.0


@pytest.mark.parametrize("g", [0.7, 0.9, 0.95, -0.5])
def test_hg_sampling(g):
    xi = (np.arange(200000) + 0.5) / 200000
    ct = np.array([lib().ymc_oracle_hg_costheta(g, float(v)) for v in xi[::50]])
    # HG CDF: F(mu) = (1-g^2)/(2g) * (1/sqrt(1+g^2-2g mu) - 1/(1+g)); the sampler maps xi to F^-1(xi)
    mu = ct.astype(np.float64)
    F = (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * mu) - 1 / (1 + g))
    assert np.max(np.abs(F - xi[::50])) < 2e-4
    ct_all = np.array([lib().ymc_oracle_hg_costheta(g, float(v)) for v in xi[::5]])
    assert ct_all.mean() == pytest.approx(g, abs=2e-3)
    assert ct_all.min() >= -1.0001 and ct_all.max() <= 1.0001


def test_specular():
    for n0, n1 in ((1.0, 1.4), (1.0, 1.44), (1.0, 1.0), (1.33, 1.0)):
        assert lib().ymc_oracle_specular(n0, n1) == pytest.approx(((n0 - n1) / (n0 + n1)) ** 2, rel=1e-6)
