"""Pins of the oracle's Smagorinsky LES branch (CollisionEgglesSomers.hpp:31-91)
and of the normal-stress moment rows (M05, M07, M10) it reads.

The strain the reference forms is, per component (quirk A.7 of the reference:
a momentum component stands where the equilibrium stress j_b u_a belongs),

    S_aa        = fct (j_a + F_a u_a / 2 - M_aa)       M_xx, M_yy, M_zz = M05, M07, M10
    2 S_ab      = 2 fct (j_b + F_b u_a / 2 - Pi_ab)    (a, b) = (x, y), (x, z), (y, z)
    fct         = 3 / (rho (1 + 6 (Nu_old + nu)))
    Nu          = cs0^2 sqrt|2 S:S - (2/3) (tr S)^2|

These tests never evaluate that expression.  They pin it through what it must
reduce to:

* at j = 0 and F = 0 it is textbook Smagorinsky on the deviatoric lattice stress
  Pi_ab = sum_d f_d e_da e_db, formed here from the populations and the direction
  table (reference Header.h:70-89), not from the oracle's moment rows;
* with an isotropic stress, each momentum / force term alone gives a hand-derived
  multiple of fct^2 J^2, distinct per component (4/3, 16/3, 28/3, ...), so a
  component read from the wrong moment changes the number;
* fct takes the previous step's Nu and the relaxation the new one;
* a transverse wave along a lattice diagonal, whose strain is a pure normal-stress
  difference, decays at the input viscosity: this pins the relaxation of M05,
  M07 and M10 the way the axis-aligned shear wave of test_oracle.py pins M06.

tests/test_oracle_mutants.py rebuilds the oracle with planted slips in this
branch and checks that this file fails on each of them.  CPU only.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

E = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1],
              [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]])
OPP = np.array([d ^ 1 for d in range(18)])
W = np.where(np.abs(E).sum(axis=1) == 1, 1.0 / 12.0, 1.0 / 24.0)   # rest populations / rho
M05, M06, M07, M08, M09, M10 = 4, 5, 6, 7, 8, 9                    # 0-based moment slots

NU, CS0 = 0.004, 0.1


def one_cell(dtype=O.F64, use_les=True, **flow):
    vals = dict(nu=NU, cs0=CS0, g3=0.8)
    vals.update(flow)
    return O.OracleCU(1, 1, 1, ghost=1, dtype=dtype, layout=O.LIJK, streaming=O.SIMPLE,
                      use_les=use_les, **vals)


def collide(cu, m, f=(0.0, 0.0, 0.0), nu_old=0.0):
    """Collide one cell holding moments m with force f and previous eddy viscosity
    nu_old; returns (populations, new Nu)."""
    cu.write_cells(np.asarray(m, dtype=np.float64).reshape(1, 1, 1, 18))
    F = np.zeros((3, 3, 3, 3))
    F[1, 1, 1] = f
    cu.set(O.FIELD_F, F.reshape(-1))
    nu = np.zeros((3, 3, 3))
    nu[1, 1, 1] = nu_old
    cu.set(O.FIELD_NU, nu.reshape(-1))
    cu.collision()
    return cu.read_cells().reshape(18), float(cu.get(O.FIELD_NU).reshape(3, 3, 3)[1, 1, 1])


def fct(rho, nu_old=0.0):
    return 3.0 / (rho * (1.0 + 6.0 * (nu_old + NU)))


def symmetric_populations(rng, rho=1.0, amp=0.02):
    """Populations with f_d = f_opp(d): zero momentum, a generic anisotropic stress."""
    eps = amp * rng.standard_normal(18)
    eps = 0.5 * (eps + eps[OPP])
    return rho * W * (1.0 + eps)


def lattice_stress(f):
    return np.einsum("d,da,db->ab", f, E, E)


def deviator(t):
    return t - np.trace(t) / 3.0 * np.eye(3)


# ---------------------------------------------------------------- (i) textbook limit
@pytest.mark.parametrize("nu_old", [0.0, 0.013])
def test_les_at_rest_is_smagorinsky_on_deviatoric_lattice_stress(nu_old):
    """j = 0, F = 0: Nu = cs0^2 fct sqrt(2 Pi_dev : Pi_dev), with Pi the lattice sum
    of the populations.  A diagonal strain read from the wrong normal-stress moment
    (M07 for M05, ...) or a cross term from the wrong shear moment breaks it."""
    rng = np.random.default_rng(11)
    for _ in range(12):
        rho = 1.0 + 0.05 * rng.standard_normal()
        f = symmetric_populations(rng, rho)
        m = O.moments_of(f)
        assert np.all(np.abs(m[1:4]) < 1e-15)
        pd = deviator(lattice_stress(f))
        exp = CS0 ** 2 * fct(f.sum(), nu_old) * math.sqrt(2.0 * np.sum(pd * pd))
        _, got = collide(one_cell(), m, nu_old=nu_old)
        assert got == pytest.approx(exp, rel=1e-11)


def test_les_at_rest_float_build_within_parity_bar():
    """The float oracle gives the same eddy viscosity to float rounding."""
    rng = np.random.default_rng(12)
    for _ in range(6):
        f = symmetric_populations(rng)
        pd = deviator(lattice_stress(f))
        exp = CS0 ** 2 * fct(f.sum()) * math.sqrt(2.0 * np.sum(pd * pd))
        _, got = collide(one_cell(O.F32), O.moments_of(f))
        assert got == pytest.approx(exp, rel=1e-5)


def test_les_at_rest_is_invariant_under_axis_permutations():
    """Permuting the axes of the populations permutes Pi and leaves Nu unchanged."""
    rng = np.random.default_rng(13)
    f = symmetric_populations(rng)
    index = {tuple(e): d for d, e in enumerate(E)}
    nus = []
    for perm in [(0, 1, 2), (1, 0, 2), (2, 1, 0), (0, 2, 1), (1, 2, 0), (2, 0, 1)]:
        g = np.empty(18)
        for d in range(18):
            g[index[tuple(E[d][list(perm)])]] = f[d]
        nus.append(collide(one_cell(), O.moments_of(g))[1])
    assert max(nus) - min(nus) <= 1e-13 * max(nus)


# ---------------------------------------------------------------- (ii) momentum terms
def isotropic_moments(rho, j, c=0.003, seed=0):
    """Moments with an isotropic stress (M05 = M07 = M10 = c, M06 = M08 = M09 = 0);
    the third- and fourth-order moments are arbitrary (the strain does not read them)."""
    m = np.random.default_rng(seed).normal(0.0, 1e-3, 18)
    m[0] = rho
    m[1:4] = j
    m[[M05, M07, M10]] = c
    m[[M06, M08, M09]] = 0.0
    return m


J, G, RHO = 0.02, 0.03, 1.1


def nu_of(sh_over_fct2, rho=RHO, nu_old=0.0):
    return CS0 ** 2 * fct(rho, nu_old) * math.sqrt(sh_over_fct2)


def two_diagonal(a, b):
    """2 sum_a S_aa^2 - (2/3)(sum_a S_aa)^2 over fct^2 for diagonal strains (a, b, 0)."""
    return 2.0 * (a * a + b * b - (a + b) ** 2 / 3.0)


# Each case: (j, F, sh / fct^2 derived by hand from the strain's stated form).
#  j = (J,0,0):  S_xx = fct J;                                   sh = 4/3 J^2
#  j = (0,J,0):  S_yy = fct J, 2S_xy = 2 fct J;                  sh = (4/3 + 4) J^2
#  j = (0,0,J):  S_zz = fct J, 2S_xz = 2S_yz = 2 fct J;          sh = (4/3 + 8) J^2
#  F = (G,0,0):  u_x = G/2rho, S_xx = fct G^2/4rho;              sh = 4/3 (G^2/4rho)^2
#  j = (J,0,0), F = (0,G,0):  u = (J/rho, G/2rho, 0);  S_xx = fct J,  S_yy = fct G^2/4rho,
#                2S_xy = 2 fct (F_y u_x / 2) = fct G J / rho
#  j = (J,0,0), F = (0,0,G):  S_xx = fct J,  S_zz = fct G^2/4rho,  2S_xz = fct G J / rho
#  j = (0,J,0), F = (0,0,G):  S_yy = fct J,  S_zz = fct G^2/4rho,  2S_xy = 2 fct J,
#                2S_yz = fct G J / rho,  2S_xz = 2 fct (F_z u_x / 2) = 0
B = G * G / (4.0 * RHO)
MOMENTUM_CASES = [
    ((J, 0, 0), (0, 0, 0), 4.0 / 3.0 * J * J),
    ((0, J, 0), (0, 0, 0), 16.0 / 3.0 * J * J),
    ((0, 0, J), (0, 0, 0), 28.0 / 3.0 * J * J),
    ((0, 0, 0), (G, 0, 0), 4.0 / 3.0 * B * B),
    ((0, 0, 0), (0, G, 0), 4.0 / 3.0 * B * B),
    ((0, 0, 0), (0, 0, G), 4.0 / 3.0 * B * B),
    ((J, 0, 0), (0, G, 0), two_diagonal(J, B) + (G * J / RHO) ** 2),
    ((J, 0, 0), (0, 0, G), two_diagonal(J, B) + (G * J / RHO) ** 2),
    ((0, J, 0), (0, 0, G), two_diagonal(J, B) + 4.0 * J * J + (G * J / RHO) ** 2),
    ((0, J, 0), (G, 0, 0), two_diagonal(J, B) + 4.0 * J * J),
    ((0, 0, J), (G, 0, 0), two_diagonal(J, B) + 8.0 * J * J),
    ((0, 0, J), (0, G, 0), two_diagonal(J, B) + 8.0 * J * J),
]


@pytest.mark.parametrize("j,F,sh", MOMENTUM_CASES)
def test_les_momentum_and_force_terms_per_component(j, F, sh):
    """Isotropic stress: the strain is carried by the momentum and force terms alone."""
    _, got = collide(one_cell(), isotropic_moments(RHO, j), F)
    assert got == pytest.approx(nu_of(sh), rel=1e-11)


def test_les_strain_does_not_read_higher_moments():
    """Third- and fourth-order moments do not enter the eddy viscosity."""
    nus = {collide(one_cell(), isotropic_moments(RHO, (J, 0, 0), seed=s))[1] for s in range(4)}
    assert len(nus) == 1


# ---------------------------------------------------------------- (ii') old and new Nu
def test_les_fct_uses_previous_step_eddy_viscosity():
    _, got = collide(one_cell(), isotropic_moments(RHO, (0, J, 0)), nu_old=0.021)
    assert got == pytest.approx(nu_of(16.0 / 3.0 * J * J, nu_old=0.021), rel=1e-11)


def test_les_relaxation_uses_the_new_eddy_viscosity():
    """Post-collision shear moments relax with b, c of nu + Nu_new (not Nu_old):
    moments(populations)[M06] / 24 ... checked through the collision's inverse."""
    rng = np.random.default_rng(17)
    f = symmetric_populations(rng)
    f[6] += 1e-3                        # a little momentum, so 2 j u terms are present
    m = O.moments_of(f)
    F = np.array([1e-4, -2e-4, 3e-4])
    nu_old = 0.017
    p, nu_new = collide(one_cell(), m, F, nu_old=nu_old)
    a = O.moments_of(p)
    rho, jv = m[0], m[1:4]
    u = (jv + 0.5 * F) / rho
    nut = nu_new + NU
    b, c = 1.0 / (1.0 + 6.0 * nut), 1.0 - 6.0 * nut
    for s, (x, y) in {M05: (0, 0), M06: (0, 1), M07: (1, 1), M08: (0, 2), M09: (1, 2), M10: (2, 2)}.items():
        assert a[s] == pytest.approx(b * (2.0 * (jv[x] + 0.5 * F[x]) * u[y] - c * m[s]), rel=1e-10, abs=1e-16)
    nut_old = nu_old + NU
    b_old, c_old = 1.0 / (1.0 + 6.0 * nut_old), 1.0 - 6.0 * nut_old
    assert abs(a[M06] - b_old * (2.0 * (jv[0] + 0.5 * F[0]) * u[1] - c_old * m[M06])) > 1e-9


# ---------------------------------------------------------------- (iii) normal-stress decay
@pytest.mark.parametrize("plane", [(0, 1), (0, 2), (1, 2)])
def test_diagonal_transverse_wave_decays_at_lattice_viscosity(plane):
    """u = U (e_a - e_b) sin(q (x_a + x_b)) on a periodic lattice: transverse to its
    wave vector (|k|^2 = 2 q^2), so it decays as exp(-nu |k|^2 t); in lattice axes its
    strain is S_aa = -S_bb with S_ab = 0, a normal-stress difference carried by
    M05/M07/M10 alone."""
    n, nu, U = 32, 0.05, 1e-5
    a, b = plane
    shape = [2, 2, 2]
    shape[a] = shape[b] = n
    cu = O.OracleCU(*shape, dtype=O.F64, layout=O.LIJK, streaming=O.ESOTWIST, nu=nu, g3=0.0)
    q = 2 * np.pi / n
    grid = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    phase = np.sin(q * (grid[a] + grid[b]))
    m = np.zeros(tuple(shape) + (18,))
    m[..., 0] = 1.0
    m[..., 1 + a] = U * phase
    m[..., 1 + b] = -U * phase
    cu.write_cells(m)

    def amp():
        return 2.0 * np.mean(cu.read_cells()[..., 1 + a] * phase)

    def step(k):
        for _ in range(k):
            cu.collision()
            cu.streaming_push()
            cu.moments()

    step(20)
    a0 = amp()
    step(100)
    a1 = amp()
    assert a1 / a0 == pytest.approx(math.exp(-nu * 2 * q * q * 100), rel=5e-3)
