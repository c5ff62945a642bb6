"""Oracle Fresnel direction update (src/mc3d_gpu.cu:226-257) against the laws of optics (CPU).

One oracle step (`ymc_oracle_step`) moves a packet from a voxel of index ni across one face into a
voxel of index nt, with its MWC state chosen so that the step's first draw (the Fresnel draw, :234)
falls above or below the reflectance R.  The new direction is then compared with physics, computed
here in binary64 from the incident direction alone:
- transmission (Snell): the components along the face scale by ni/nt, the normal component becomes
  cos(theta_t) = sqrt(1 - (ni/nt)^2 sin^2(theta_i)) with the sign of travel, so
  ni sin(theta_i) = nt sin(theta_t) and the direction stays a unit vector;
- reflection: only the component normal to the crossed face changes sign;
- total internal reflection: beyond the critical angle every draw reflects.
Every crossing axis (x, y, z) is taken in both directions of travel.  The reference's rule for a
packet that hits an edge (the x face counts before y, y before z, :227) is a choice of the
reference, not physics; it is checked as that rule.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

M32 = (1 << 32) - 1
A_STREAM = 4294966893  # an odd multiplier of the table (a*2^32-1 and a*2^31-1 prime)
VOX = 0.01


def _state_with_first_draw(target_co: float, a: int = A_STREAM) -> int:
    """MWC state whose next draw rand_MWC_co (gpumcml_rng.cu:38-46) is about target_co."""
    lo_target = min(int(target_co * 2 ** 32), M32 - 1)
    inv = pow(a, -1, 1 << 32)
    carry = 12345
    lo = ((lo_target - carry) * inv) % (1 << 32)
    x = (carry << 32) | lo
    nxt = (lo * a + carry) & M32
    assert nxt == lo_target
    return x


def _scene(ni: float, nt: float, axis: int, sign: int):
    """3x3x3 grid of 0.01 cm voxels: row 1 (index ni) everywhere, row 2 (index nt) in the voxel beside
    the centre along `axis` in direction `sign`.  Non-absorbing rows so that only the crossing acts."""
    rows = np.ones((3, 3, 3), np.uint8)
    c = [1, 1, 1]
    c[axis] += sign
    rows[c[2], c[1], c[0]] = 2
    table = (np.array([0, 0.5, 0.5], np.float32), np.array([0, 50, 50], np.float32),
             np.array([0, 0.9, 0.9], np.float32), np.array([1.0, ni, nt], np.float32))
    return rows, table


def _direction(axis: int, sign: int, theta: float, psi: float):
    """Unit vector at angle theta to the normal of the face crossed along `axis` in direction `sign`;
    psi turns the tangential part about the normal."""
    t1, t2 = [a for a in range(3) if a != axis]
    u = [0.0, 0.0, 0.0]
    u[axis] = sign * math.cos(theta)
    u[t1] = math.sin(theta) * math.cos(psi)
    u[t2] = math.sin(theta) * math.sin(psi)
    u32 = np.array(u, np.float32)
    return u32 / np.float32(np.sqrt(np.sum(u32.astype(np.float64) ** 2)))


def _cross(ni, nt, axis, sign, theta, psi, draw, mode=0):
    rows, table = _scene(ni, nt, axis, sign)
    u = _direction(axis, sign, theta, psi)
    # 1e-4 cm before the crossed face and 0.005 cm from the others: up to theta = 1.5 the flight
    # meets the crossed face first
    pos = [0.015, 0.015, 0.015]
    pos[axis] = 0.015 + sign * 0.0049
    state = (*pos, *u, 1.0, 1e3)  # s = 1e3: the flight reaches the face
    frozen, st, ix, _, fres = O.step(rows, (VOX, VOX, VOX), table, state, (1, 1, 1, 1), _state_with_first_draw(draw),
                                     A_STREAM, fresnel_mode=mode)
    assert not frozen and fres == 1
    # the packet stopped on the crossed face
    assert st[axis] == np.float32((1 + (sign > 0)) * VOX)
    return u.astype(np.float64), st[3:6].astype(np.float64), ix


def _snell(u, axis, ni, nt):
    eta = ni / nt
    cos_i = abs(u[axis])
    sin_t2 = eta * eta * (1 - cos_i * cos_i)
    t = eta * u.copy()
    t[axis] = math.copysign(math.sqrt(1 - sin_t2), u[axis])
    return t


PAIRS = [(1.4, 1.0), (1.0, 1.4), (1.39, 1.44), (1.44, 1.39), (1.33, 1.6)]
CASES = [(axis, sign) for axis in range(3) for sign in (1, -1)]


@pytest.mark.parametrize("axis,sign", CASES)
@pytest.mark.parametrize("ni,nt", PAIRS)
def test_transmission_obeys_snell(axis, sign, ni, nt):
    crit = math.asin(nt / ni) if ni > nt else math.pi / 2
    for theta in (0.1, 0.4, 0.7, min(1.2, crit - 0.05)):
        u, new, ix = _cross(ni, nt, axis, sign, theta, 0.6, draw=0.999)
        want = _snell(u, axis, ni, nt)
        assert np.allclose(new, want, atol=2e-6), (theta, new, want)
        assert abs(np.linalg.norm(new) - 1) < 1e-6
        # ni sin(theta_i) = nt sin(theta_t), normal component keeps the direction of travel
        sin_i = math.sqrt(max(0.0, 1 - u[axis] ** 2))
        sin_t = math.sqrt(max(0.0, 1 - new[axis] ** 2))
        assert abs(ni * sin_i - nt * sin_t) < 4e-6
        assert np.sign(new[axis]) == sign
        # the packet entered the neighbour voxel (row 2)
        want_ix = [1, 1, 1]
        want_ix[axis] += sign
        assert ix.tolist() == want_ix + [2]
        # an inverted ratio nt/ni would move the tangential part the other way
        wrong = _snell(u, axis, nt, ni) if (ni / nt) ** 2 * (1 - u[axis] ** 2) <= 1 and \
            (nt / ni) ** 2 * (1 - u[axis] ** 2) <= 1 else None
        if wrong is not None:
            assert np.max(np.abs(wrong - new)) > 1e-3


@pytest.mark.parametrize("axis,sign", CASES)
@pytest.mark.parametrize("ni,nt", PAIRS)
def test_reflection_flips_only_the_normal_component(axis, sign, ni, nt):
    for theta in (0.2, 0.6, 1.0):
        u, new, ix = _cross(ni, nt, axis, sign, theta, 1.1, draw=0.0)
        want = u.copy()
        want[axis] = -want[axis]
        assert np.allclose(new, want, atol=1e-6), (theta, new, want)
        # a reflected packet stays in its voxel (src/mc3d_gpu.cu:334-339)
        assert ix.tolist() == [1, 1, 1, 1]


@pytest.mark.parametrize("axis,sign", CASES)
@pytest.mark.parametrize("mode", [0, 1])
def test_total_internal_reflection_beyond_the_critical_angle(axis, sign, mode):
    ni, nt = 1.44, 1.0
    crit = math.asin(nt / ni)
    for theta in (crit + 0.02, crit + 0.3, 1.5):
        u, new, ix = _cross(ni, nt, axis, sign, theta, 0.3, draw=0.9999, mode=mode)
        want = u.copy()
        want[axis] = -want[axis]
        assert np.allclose(new, want, atol=1e-6)
        assert ix.tolist() == [1, 1, 1, 1]
    # just below the critical angle the same draw transmits
    u, new, ix = _cross(ni, nt, axis, sign, crit - 0.02, 0.3, draw=0.9999, mode=mode)
    assert np.sign(new[axis]) == sign and ix[3] == 2


@pytest.mark.parametrize("ni,nt", [(1.4, 1.0), (1.33, 1.6)])
def test_reflect_or_transmit_follows_the_draw_against_R(ni, nt):
    """The first draw decides: eps < R reflects (:234-249).  R at the crossing angle is the exact
    Fresnel reflectance in mode 1."""
    theta = 0.5
    cos_i = math.cos(theta)
    sin_t = ni / nt * math.sin(theta)
    cos_t = math.sqrt(1 - sin_t ** 2)
    rs = ((ni * cos_i - nt * cos_t) / (ni * cos_i + nt * cos_t)) ** 2
    rp = ((ni * cos_t - nt * cos_i) / (ni * cos_t + nt * cos_i)) ** 2
    R = 0.5 * (rs + rp)
    for draw, reflects in ((R * 0.9, True), (R * 1.1, False)):
        u, new, ix = _cross(ni, nt, 0, 1, theta, 0.0, draw=draw, mode=1)
        assert (new[0] < 0) == reflects
        assert (ix[3] == 1) == reflects


def _edge_start(ux: np.float32, uz: np.float32, face: np.float32):
    """Start coordinates x0, z0 below `face` from which the x and the z face are met at the same
    binary32 distance t = (face - x0) / ux = (face - z0) / uz (the expression of :31-36)."""
    z0 = np.float32(face - np.float32(0.1) * uz)
    for _ in range(256):
        tz = np.float32(np.float32(face - z0) / uz)
        x_mid = np.float32(face - ux * tz)
        xs = x_mid + np.arange(-64, 65, dtype=np.float32) * np.spacing(x_mid)
        tx = (np.float32(face) - xs) / ux
        hit = np.nonzero(tx == tz)[0]
        if hit.size:
            return xs[hit[0]], z0
        z0 = np.nextafter(z0, np.float32(0))
    raise AssertionError("no start point found")


def test_edge_hit_takes_the_x_face_first():
    """A flight that reaches the x and z faces at once (an edge) crosses by the x face's normal: the
    reference tests i before j before k (:227).  The x and z direction components differ, so the
    crossing axis decides the outcome; index 1.0 beyond both faces."""
    rows = np.ones((3, 3, 3), np.uint8)
    rows[2, 1, 2] = 2
    rows[1, 1, 2] = 2
    rows[2, 1, 1] = 2
    table = (np.array([0, 0.5, 0.5], np.float32), np.array([0, 50, 50], np.float32),
             np.array([0, 0.9, 0.9], np.float32), np.array([1.0, 1.4, 1.0], np.float32))
    u = np.array([math.cos(0.6), 0, math.sin(0.6)], np.float32)
    face = np.float32(0.5)
    x0, z0 = _edge_start(u[0], u[2], face)
    state = (x0, 0.375, z0, *u, 1.0, 1e3)
    frozen, st, ix, _, fres = O.step(rows, (0.25, 0.25, 0.25), table, state, (1, 1, 1, 1),
                                     _state_with_first_draw(0.999), A_STREAM)
    assert fres == 1 and st[0] == face and st[2] == face
    assert ix.tolist() == [2, 1, 2, 2]
    ud = u.astype(np.float64)
    want = _snell(ud, 0, 1.4, 1.0)
    assert np.allclose(st[3:6], want / np.linalg.norm(want), atol=2e-6)
    # by the z face's normal the same flight is beyond the critical angle and would reflect
    assert (1.4 / 1.0) ** 2 * (1 - ud[2] ** 2) > 1
    assert st[5] > 0 and abs(st[5] - ud[2]) > 0.1
