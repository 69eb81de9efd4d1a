"""The boundary 1-D model of Dirichlet inimod 2/3: p_pfields, the Wait
recursion and ezl (boundary_conds.f90:436-598) as the oracle restates them.

The reference build pins the whole lift (RHS and readout) on the parity cases;
the closed forms below pin the pieces by routes that do not retype the
reference's algebra: the half-space field solves E'' = i alpha E below the
surface and is linear in the air, the Wait recursion agrees with the tanh form
of the layered-earth impedance, and ezl's layer-1 expression reduces to 1 at
the surface (field continuity).
"""
import dataclasses

import numpy as np
import pytest

from movfem_amd import model as M
from oracle.oracle import Oracle, ref_available
from tests.parity_cases import cases

MU0 = 4 * M.PI * 1e-7
B0 = 1e-9


def _model(inimod, **kw):
    m = M.random_small(8, 2, 2, 3, seed=11, nz_air=1, freqs=(0.3, 40.0))
    return dataclasses.replace(m, inimod=inimod, **kw)


def _gamma(w, sig):
    ps = M.f32(sig) + 1j * M.f32(M.EPS0 * w)
    return np.sqrt(1j * w * MU0 * ps)


@pytest.mark.parametrize("fi", [0, 1])
def test_halfspace_field_ode(fi):
    """inimod 2: below the surface fp'' = (i alpha) fp with fp(0) = 1/gamma; in
    the air fp is linear with fp(0) = 1/gamma_air and slope -1 (z in km)."""
    m = _model(2, h_sigma=0.03)
    o = Oracle(m)
    o.set_frequency(fi)
    w = m.omega(fi)
    c = M.f32(2 * w * B0)
    ge, ga = _gamma(w, 0.03), _gamma(w, 0.0)
    fp = lambda zn: o.bd_field(zn)[1] / (1j * c)  # polarization 2 carries +i c fp
    # polarizations differ by sign only
    for zn in (m.ztop - 700.0, m.ztop + 150.0):
        e = o.bd_field(zn)
        assert abs(e[0] + e[1]) <= 1e-15 * abs(e[1])
    # surface values
    assert abs(fp(m.ztop) - 1 / ge) <= 1e-13 * abs(1 / ge)
    assert abs(fp(m.ztop + 1e-9) - 1 / ga) <= 1e-9 * abs(1 / ga)
    # ODE in depth (km): second difference ~ gamma^2 fp (step of ~0.01 skin
    # depths: truncation ~1e-5, rounding ~1e-12)
    h = 10.0 / abs(ge)  # metres: 0.01 / |gamma| in the km depth unit
    zn = m.ztop - 4 * h
    d2 = (fp(zn - h) - 2 * fp(zn) + fp(zn + h)) / (h / 1000.0) ** 2
    assert abs(d2 - ge ** 2 * fp(zn)) <= 1e-4 * abs(ge ** 2 * fp(zn))
    # air: exactly linear, slope d fp / d z_km = -1 with z = (ztop - zn)/1000
    z1, z2 = m.ztop + 100.0, m.ztop + 1.0e6
    slope = (fp(z2) - fp(z1)) / ((m.ztop - z2) / 1000.0 - (m.ztop - z1) / 1000.0)
    assert abs(slope + 1.0) <= 1e-9


def _tanh_cz1(w, sig, dz):
    """Layered-earth impedance by the tanh recursion (Wait 1954):
    C_n = 1/k_n, C_l = (k_l C_{l+1} + tanh(k_l d_l)) / (k_l (1 + k_l C_{l+1} tanh(k_l d_l)))."""
    k = [_gamma(w, s) for s in sig]
    cz = 1 / k[-1]
    for l in range(len(sig) - 2, -1, -1):
        t = np.tanh(k[l] * dz[l])
        cz = (k[l] * cz + t) / (k[l] * (1 + k[l] * cz * t))
    return cz


@pytest.mark.parametrize("layers", [((0.05,), ()), ((0.05, 0.002), (0.6,)),
                                    ((0.02, 0.0015, 0.4), (0.35, 0.9)), ((0.3, 0.3, 0.3), (0.2, 5.0))])
@pytest.mark.parametrize("fi", [0, 1])
def test_layered_impedance_and_continuity(layers, fi):
    """inimod 3: the air field is i c (cz1 - z) with cz1 the tanh-form impedance
    (1/gamma for a uniform stack); at the surface node (z = 0, earth branch)
    ezl = 1, so the earth expression meets the air one; below the surface the
    reference's ezl layer test (z >= 0 against interfaces zl <= 0) gives 0."""
    sig, dz = layers
    m = _model(3, l_sigma=sig, l_dz=dz)
    o = Oracle(m)
    o.set_frequency(fi)
    w = m.omega(fi)
    c = M.f32(2 * w * B0)
    cz1 = _tanh_cz1(w, sig, dz)
    if len(set(sig)) == 1:
        assert abs(cz1 - 1 / _gamma(w, sig[0])) <= 1e-12 * abs(cz1)
    zn = m.ztop + 250.0
    z = (m.ztop - zn) / 1000.0
    e = o.bd_field(zn)
    assert abs(e[0] - 1j * c * (cz1 - z)) <= 1e-12 * abs(cz1 - z) * c
    assert abs(e[1] + e[0]) <= 1e-15 * abs(e[0])
    es = o.bd_field(m.ztop)
    assert abs(es[0] - 1j * c * cz1) <= 1e-13 * abs(c * cz1)
    assert np.all(o.bd_field(m.ztop - 10.0) == 0)


def test_inimod1_values_zero():
    m = _model(1)
    o = Oracle(m)
    o.set_frequency(0)
    assert np.all(o.dirichlet_values(1, 1, 1) == 0)


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("bd", ["halfspace", "layers3", "layer1"])
@pytest.mark.parametrize("name", sorted(n for n, m in cases().items() if m.dirichlet))
def test_lift_against_reference(name, bd):
    """global_vfem's RHS with the lift (MoVFEM_3DMT.f90:251-261) and
    node_solution with the Dirichlet edge values, against the reference's
    modules; bit-exact where the element RHS is (8- and 20-node)."""
    from oracle.refrun import run_reference

    kw = {"halfspace": dict(inimod=2, h_sigma=0.02),
          "layers3": dict(inimod=3, l_sigma=(0.05, 0.002, 0.3), l_dz=(0.4, 1.1)),
          "layer1": dict(inimod=3, l_sigma=(0.03,), l_dz=())}[bd]
    m = dataclasses.replace(cases()[name], **kw)
    o = Oracle(m)
    x = np.array([1, 1j]) @ np.random.Generator(np.random.PCG64(5)).standard_normal((2, 2 * o.nne))
    r = run_reference(m, vfem_freqs=[1, 2], readout_x=x)
    exact = m.nodes_per_elem != 27
    for fi in (1, 2):
        o.set_frequency(fi - 1)
        _, _, _, b = o.global_vfem()
        br = r[f"b_{fi}"]
        if exact:
            assert np.array_equal(b, br)
        else:
            assert np.abs(b - br).max() <= 1e-13 * np.abs(br).max()
    o.set_frequency(1)
    e, h = o.node_solution(x)
    assert np.array_equal(e, r["esol"]) and np.array_equal(h, r["hsol"])
