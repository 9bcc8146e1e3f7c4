"""Pins of the oracle's general-cell (triclinic) nonbonded sum,
oracle_nb_fp64_cell (brute force over pairs and the 27 images around the
fractional minimum image), by what periodicity and geometry fix:
  * an orthorhombic cell reproduces the cell-list oracle oracle_nb_fp64;
  * a unimodular change of lattice basis (the same lattice) changes nothing;
  * a rigid rotation of positions and cell leaves energies and the virial's
    trace unchanged and rotates forces;
  * forces are minus the energy gradient (central differences) in a skewed
    cell.
"""
import math

import numpy as np
import pytest

from mi355nb import systems
from oracle import oracle as orc

COULOMB = 332.0636
CELL = np.array([[26.0, 0.0, 0.0], [7.0, 25.0, 0.0], [-5.0, 6.0, 26.5]])


def _run(s, pos=None, cell=None):
    p = s.params
    A, B = systems.lj_tables(s)
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    return orc.nb_fp64_cell(s.pos if pos is None else pos, s.charge * math.sqrt(COULOMB), s.vdw_type, A, B,
                            s.excl_ptr, s.excl_idx, s.cell if cell is None else cell, p.cutoff, p.switchdist, beta)


@pytest.fixture(scope="module")
def tri():
    base = systems.small_box(n_waters=300, n_chains=2, chain_len=20, box=26.0, seed=12)
    return base, systems.sheared(base, CELL)


def test_orthorhombic_cell_matches_cell_list_oracle(tri):
    base, _ = tri
    p = base.params
    A, B = systems.lj_tables(base)
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    q = base.charge * math.sqrt(COULOMB)
    o = orc.nb_fp64(base.pos, q, base.vdw_type, A, B, base.excl_ptr, base.excl_idx, base.box, p.cutoff,
                    p.switchdist, beta)
    c = _run(base, cell=np.diag(base.box))
    assert c["npairs"] == o["npairs"]
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(c[k] - o[k]) <= 1e-11 * abs(o[k]), k
    assert np.abs(c["f"] - o["f"]).max() <= 1e-9 * np.abs(o["f"]).max()


def test_change_of_lattice_basis(tri):
    _, s = tri
    o = _run(s)
    a, b, c = s.cell
    other = np.array([a, b + a, c - b + a])        # unimodular: the same lattice
    assert abs(np.linalg.det(other) - np.linalg.det(s.cell)) < 1e-9
    r = _run(s, cell=other)
    assert r["npairs"] == o["npairs"]
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(r[k] - o[k]) <= 1e-11 * abs(o[k]), k
    assert np.abs(r["f"] - o["f"]).max() <= 1e-9 * np.abs(o["f"]).max()


def test_rigid_rotation(tri):
    _, s = tri
    o = _run(s)
    th = 0.7
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, math.cos(0.3), -math.sin(0.3)], [0, math.sin(0.3), math.cos(0.3)]])
    Rm = Rx @ Rz
    r = _run(s, pos=s.pos @ Rm.T, cell=s.cell @ Rm.T)
    assert r["npairs"] == o["npairs"]
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(r[k] - o[k]) <= 1e-10 * abs(o[k]), k
    assert np.abs(r["f"] - o["f"] @ Rm.T).max() <= 1e-8 * np.abs(o["f"]).max()
    assert abs(np.trace(r["virial"]) - np.trace(o["virial"])) <= 1e-9 * abs(np.trace(o["virial"]))


def test_forces_are_energy_gradient(tri):
    _, s = tri
    o = _run(s)
    rng = np.random.default_rng(1)
    for a in rng.choice(s.natoms, 5, replace=False):
        for d in range(3):
            h = 1e-5
            xp = s.pos.copy(); xp[a, d] += h
            xm = s.pos.copy(); xm[a, d] -= h
            ep, em = _run(s, pos=xp), _run(s, pos=xm)
            g = sum(ep[k] - em[k] for k in ("e_vdw", "e_elec", "e_slow")) / (2 * h)
            f = o["f"][a, d] + o["fs"][a, d]
            assert abs(f + g) <= 2e-5 * max(1.0, abs(f)), (a, d, f, -g)


def test_orthorhombic_cell_matches_cell_list_oracle_force_switched(tri):
    base, _ = tri
    p = base.params
    A, B = systems.lj_tables(base)
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    q = base.charge * math.sqrt(COULOMB)
    o = orc.nb_fp64(base.pos, q, base.vdw_type, A, B, base.excl_ptr, base.excl_idx, base.box, p.cutoff,
                    p.switchdist, beta, vdw_force_switching=True)
    c = orc.nb_fp64_cell(base.pos, q, base.vdw_type, A, B, base.excl_ptr, base.excl_idx, np.diag(base.box), p.cutoff,
                         p.switchdist, beta, vdw_force_switching=True)
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(c[k] - o[k]) <= 1e-11 * abs(o[k]), k
    assert np.abs(c["f"] - o["f"]).max() <= 1e-9 * np.abs(o["f"]).max()
