"""NAMD's dumpbench format (src/DumpBench.C:26-311): reading a hand-written
file in that layout (tests/golden/dumpbench_seven_atoms.txt: exclusion rows
in both of NAMD's forms, the shared all-full array that flags the atom
itself and an individual array), writing the build's systems and reading
them back (write -> read -> write is byte-identical; arrays equal to the
printed precision), and the FP64 oracle on a read-back system agreeing with
the oracle on the original to the positions' printed precision."""
import io
import math
from pathlib import Path

import numpy as np
import pytest

from mi355nb import dumpbench as D
from mi355nb import systems
from oracle import oracle as orc

GOLDEN = Path(__file__).parent / "golden" / "dumpbench_seven_atoms.txt"


def _pairs(s):
    i = np.repeat(np.arange(s.natoms), np.diff(s.excl_ptr))
    return {(int(min(a, b)), int(max(a, b))) for a, b in zip(i, s.excl_idx)}


def test_golden_file_parses():
    s, info = D.read(GOLDEN)
    assert s.natoms == 7
    assert np.array_equal(s.box, [20.0, 20.0, 20.0])
    assert (s.params.cutoff, s.params.switchdist, s.params.pairlistdist) == (8.0, 7.0, 9.0)
    # the shared all-full rows flag the atom itself; it is not an exclusion
    assert _pairs(s) == {(0, 1), (0, 2), (1, 2), (0, 6), (3, 4), (3, 5), (4, 5)}
    # both directions are listed
    assert sorted(s.excl_idx[s.excl_ptr[1]:s.excl_ptr[2]].tolist()) == [0, 2]
    assert np.array_equal(s.vdw_type, [0, 0, 0, 1, 2, 2, 0])
    A, B = systems.lj_tables(s)
    assert A[0, 1] == A[1, 0] == 500000.0 and B[1, 1] == 595.0 and A[2, 2] == 0.0
    assert np.allclose(s.pos[6], [4.0, 9.0, 5.0]) and s.charge[3] == -0.834
    assert [p.tolist() for p in info["patches"]] == [[0, 1, 2, 6], [3, 4, 5]]
    assert info["computes"] == [(0, 13, 1, 13)]
    assert D.lattice_index(0, 0, 0) == 13 and D.lattice_index(1, -1, 0) == 11
    assert D.lattice_index(-1, -1, -1) == 0 and D.lattice_index(1, 1, 1) == 26


def test_rejects_what_the_path_does_not_compute():
    text = GOLDEN.read_text()
    with pytest.raises(ValueError, match="right-handed"):
        D.read(text.replace("cellBasisVector2 0.000000 20.000000", "cellBasisVector2 0.000000 -20.000000"))
    # a general cell is read as System.cell
    sk, _ = D.read(text.replace("cellBasisVector2 0.000000 20.000000", "cellBasisVector2 1.000000 20.000000"))
    assert sk.cell is not None and sk.cell[1, 0] == 1.0
    # a 1-4 (modified) exclusion flag needs exclude scaled1-4
    modified = text.replace("0 1 6 1 1 0 0 0 1", "0 1 6 1 1 2 0 0 1").replace("1 4 5 1 1", "1 0 5 2 0 0 0 1 1")
    with pytest.raises(ValueError, match="modified"):
        D.read(modified)
    ok, info = D.read(modified.replace("exclude 2", "exclude 4"))
    assert info["modified_pairs"] == 1 and (0, 3) not in _pairs(ok)
    # the modified pair comes back from both ends, for set_modified_exclusions
    mp, mi = info["mod_ptr"], info["mod_idx"]
    got = {(a, int(b)) for a in range(ok.natoms) for b in mi[mp[a]:mp[a + 1]]}
    assert got == {(0, 3), (3, 0)}
    with pytest.raises(ValueError, match="missing"):
        D.read(text.replace("4.000000 9.000000 5.000000 1.000000 6 1 0 0 0 0\n", "").replace("PATCH_BEGIN\n4\n",
                                                                                               "PATCH_BEGIN\n3\n"))
    with pytest.raises(ValueError, match="fixed"):
        D.read(text.replace("5.000000 1.000000 6 1 0 0 0 0", "5.000000 1.000000 6 1 0 1 0 0"))


def test_write_read_round_trip():
    s = systems.small_box(n_waters=216, n_chains=2, chain_len=20)
    s.partition = np.zeros(s.natoms, np.int8)
    s.partition[:20] = 1
    buf = io.StringIO()
    D.write(buf, s, alch=dict(alchLambda=0.3, alchLambda2=0.35, alchVdwShiftCoeff=5.0))
    text = buf.getvalue()
    r, info = D.read(text)
    assert r.natoms == s.natoms and np.array_equal(r.box, s.box)
    assert np.abs(r.pos - s.pos).max() <= 5e-7 + 1e-12
    assert np.abs(r.charge - s.charge).max() <= 5e-7
    assert np.array_equal(r.vdw_type, s.vdw_type)
    assert _pairs(r) == _pairs(s)
    assert np.array_equal(r.partition, s.partition)
    A, B = systems.lj_tables(s)
    rA, rB = systems.lj_tables(r)
    assert np.allclose(rA, A, rtol=5e-6) and np.allclose(rB, B, rtol=5e-6)
    assert info["params"]["alchVdwShiftCoeff"] == 5.0 and info["params"]["alchOn"] == 1.0
    # the patch list covers every atom once; 13 pair computes per patch
    assert sorted(np.concatenate(info["patches"]).tolist()) == list(range(s.natoms))
    assert len(info["computes"]) == 13 * len(info["patches"])
    # write -> read -> write is byte-identical
    buf2 = io.StringIO()
    D.write(buf2, r, alch=dict(alchLambda=0.3, alchLambda2=0.35, alchVdwShiftCoeff=5.0))
    assert buf2.getvalue() == text


def test_oracle_consumes_the_file():
    s = systems.small_box(n_waters=216, n_chains=2, chain_len=20)
    buf = io.StringIO()
    D.write(buf, s)
    r, _ = D.read(buf.getvalue())
    p = s.params
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    out = []
    for x in (s, r):
        A, B = systems.lj_tables(x)
        out.append(orc.nb_fp64(x.pos, x.charge * math.sqrt(332.0636), x.vdw_type, A, B, x.excl_ptr, x.excl_idx,
                               x.box, p.cutoff, p.switchdist, beta))
    f0, f1 = out[0]["f"], out[1]["f"]
    err = float(np.sqrt(((f1 - f0) ** 2).sum() / (f0 ** 2).sum()))
    # positions rounded to 1e-6 A and LJ tables to 6 significant digits
    assert err < 1e-4, err
    assert abs(out[1]["e_vdw"] - out[0]["e_vdw"]) <= 1e-5 * abs(out[0]["e_vdw"])


def test_modified_exclusions_round_trip():
    """Scaled 1-4 exclusions written as EXCHCK_MOD flags with exclude
    scaled1-4, scale14 and the 1-4 LJ columns, read back as the pair lists
    and tables the modified-exclusion correction takes; the oracle's
    correction on the read data matches the original within the file's
    rounding (positions 1e-6 A, tables 6 significant digits)."""
    s = systems.small_box(n_waters=100, n_chains=2, chain_len=20, box=18.0, seed=4)
    mp, mi = systems.chain_pairs_14(s)
    A14, B14 = systems.lj14_tables(s)
    mod = dict(mod_ptr=mp, mod_idx=mi, scale14=0.8333, A14=A14, B14=B14)
    buf = io.StringIO()
    D.write(buf, s, modified=mod)
    text = buf.getvalue()
    r, info = D.read(text)
    assert info["params"]["exclude"] == D.EXCLUDE_SCALED14 and info["scale14"] == 0.8333
    assert info["modified_pairs"] == mi.size and _pairs(r) == _pairs(s)
    want = {(a, int(b)) for a in range(s.natoms) for b in mi[mp[a]:mp[a + 1]]}
    want |= {(b, a) for a, b in want}
    rp, ri = info["mod_ptr"], info["mod_idx"]
    assert {(a, int(b)) for a in range(r.natoms) for b in ri[rp[a]:rp[a + 1]]} == want
    assert np.allclose(info["A14"], A14, rtol=5e-6) and np.allclose(info["B14"], B14, rtol=5e-6)
    buf2 = io.StringIO()
    D.write(buf2, r, modified=dict(mod_ptr=rp, mod_idx=ri, scale14=info["scale14"], A14=info["A14"],
                                   B14=info["B14"]))
    assert buf2.getvalue() == text
    p = s.params
    out = []
    for x, (pp, ii, a14, b14) in ((s, (mp, mi, A14, B14)), (r, (rp, ri, info["A14"], info["B14"]))):
        A, B = systems.lj_tables(x)
        out.append(orc.mod_excl_fp64(x.pos, x.charge * math.sqrt(332.0636), x.vdw_type, A, B, a14, b14, pp, ii,
                                     x.box, p.cutoff, p.switchdist, 0.8333))
    assert out[0]["npairs"] == out[1]["npairs"] == mi.size
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(out[1][k] - out[0][k]) <= 1e-4 * abs(out[0][k]), k


def test_general_cell_round_trip():
    """A system in a general cell written and read back: the lattice rows,
    the patch list (binned in fractional coordinates) and a byte-identical
    rewrite; the general-cell oracle gives the same energies on the read data
    within the file's rounding."""
    base = systems.small_box(n_waters=300, n_chains=2, chain_len=20, box=30.0, seed=6)
    s = systems.sheared(base, [[30.0, 0.0, 0.0], [6.0, 29.0, 0.0], [-4.0, 5.0, 30.5]])
    buf = io.StringIO()
    D.write(buf, s)
    text = buf.getvalue()
    r, info = D.read(text)
    assert r.cell is not None and np.allclose(r.cell, s.cell, atol=1e-6)
    assert sorted(np.concatenate(info["patches"]).tolist()) == list(range(s.natoms))
    buf2 = io.StringIO()
    D.write(buf2, r)
    assert buf2.getvalue() == text
    p = s.params
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    out = []
    for x in (s, r):
        A, B = systems.lj_tables(x)
        out.append(orc.nb_fp64_cell(x.pos, x.charge * math.sqrt(332.0636), x.vdw_type, A, B, x.excl_ptr,
                                    x.excl_idx, x.cell, p.cutoff, p.switchdist, beta))
    assert out[0]["npairs"] == out[1]["npairs"]
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(out[1][k] - out[0][k]) <= 1e-4 * abs(out[0][k]), k
