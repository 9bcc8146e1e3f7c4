"""ORACLE - test infrastructure only (the checker).

ctypes wrappers over liboracle.so (nb_oracle.c).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built: run python oracle/build.py")
        _LIB = C.CDLL(path)
        L = _LIB
        L.oracle_nb_fp64.argtypes = [C.c_int, dp, dp, i32p, C.c_int, dp, dp, i64p, i32p, dp, C.c_double, C.c_double,
                                     C.c_double, dp, dp, dp, dp, dp, i64p, C.c_int]
        L.oracle_nb_fp64_vdw_force_switch.argtypes = L.oracle_nb_fp64.argtypes
        L.oracle_pairs_fp32.argtypes = [C.c_int, dp, i64p, i32p, dp, C.c_float, C.c_float, i64p, u64p, u64p,
                                        C.c_void_p, C.c_int64]
        L.oracle_pair_hash.restype = C.c_uint64
        L.oracle_pair_hash.argtypes = [C.c_int64, C.c_int64, C.c_int64]
        L.oracle_pme_spread_gpu.argtypes = [C.c_int, fp, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.oracle_pme_gather_gpu.argtypes = [C.c_int, fp, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_void_p, dp]
        L.oracle_pme_scale_coordinates.argtypes = [C.c_int, dp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.oracle_pme_spread_cpu.argtypes = [C.c_int, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.oracle_pme_gather_cpu.argtypes = [C.c_int, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
        L.oracle_bspline_gpu.argtypes = [C.c_double, C.c_int, dp, dp]
        L.oracle_pairs_exact.argtypes = [C.c_int, dp, i64p, i32p, dp, C.c_double, C.c_double, i64p, u64p, u64p,
                                         C.c_void_p, C.c_int64]
        L.oracle_namd_tables.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int),
                                         C.c_void_p, C.c_void_p]
        L.oracle_namd_cpu.argtypes = L.oracle_nb_fp64.argtypes
        L.oracle_nb_table.argtypes = [C.c_int, dp, dp, i32p, C.c_int, dp, dp, i64p, i32p, dp, C.c_double,
                                      C.c_double, C.c_int, fp, fp, dp, dp, dp, dp, dp, i64p, C.c_int]
        L.oracle_namd_gpu_tables.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, fp, fp]
        L.oracle_nb_adjoint_fp64.argtypes = [C.c_int, dp, dp, i32p, C.c_int, i64p, i32p, dp, C.c_double,
                                             C.c_double, C.c_double, dp, dp, dp, dp, C.c_int]
        i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
        L.oracle_nb_fp64_alch.argtypes = [C.c_int, dp, dp, i32p, C.c_int, dp, dp, i64p, i32p, dp, C.c_double,
                                          C.c_double, C.c_double, i8p, dp, C.c_int, dp, dp, dp, dp, dp, i64p,
                                          C.c_int]
        L.oracle_nb_fp64_alch_vdw_force_switch.argtypes = L.oracle_nb_fp64_alch.argtypes
        L.oracle_mod_excl_fp64.argtypes = [C.c_int, dp, dp, i32p, C.c_int, dp, dp, dp, dp, i64p, i32p, dp,
                                           C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, dp, dp, dp, i64p]
        L.oracle_nb_fp64_cell.argtypes = [C.c_int, dp, dp, i32p, C.c_int, dp, dp, i64p, i32p, dp, C.c_double,
                                          C.c_double, C.c_double, C.c_int, dp, dp, dp, dp, dp, i64p, C.c_int]
        L.oracle_mod_excl_cell_fp64.argtypes = L.oracle_mod_excl_fp64.argtypes
        L.oracle_mod_excl_alch_fp64.argtypes = [C.c_int, dp, dp, i32p, C.c_int, dp, dp, dp, dp, i64p, i32p, dp,
                                                C.c_double, C.c_double, C.c_double, i8p, dp, C.c_int, dp, dp, dp,
                                                dp, dp, i64p]
    return _LIB


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def nb_fp64(pos, q_scaled, vdw_type, A, B, excl_ptr, excl_idx, box, cutoff, switchdist, beta, nthreads=0,
            vdw_force_switching=False):
    """FP64 analytic oracle. Returns dict(f, fs, e_vdw, e_elec, e_slow, virial, virial_slow, npairs).
    vdw_force_switching: NAMD's vdwForceSwitching (Steinbach-Brooks force
    switch, ComputeNonbondedUtil.C:523-533, :900-919) instead of energy
    switching."""
    n = len(q_scaled)
    ntypes = int(round(np.sqrt(np.asarray(A).size)))
    f = np.zeros(3 * n)
    fs = np.zeros(3 * n)
    en = np.zeros(3)
    vir = np.zeros(9)
    virs = np.zeros(9)
    npairs = np.zeros(1, np.int64)
    fn = lib().oracle_nb_fp64_vdw_force_switch if vdw_force_switching else lib().oracle_nb_fp64
    fn(n, _c(pos, np.float64).reshape(-1), _c(q_scaled, np.float64), _c(vdw_type, np.int32),
       ntypes, _c(A, np.float64).reshape(-1), _c(B, np.float64).reshape(-1),
       _c(excl_ptr, np.int64), _c(excl_idx, np.int32), _c(box, np.float64), cutoff, switchdist,
       beta, f, fs, en, vir, virs, npairs, int(nthreads))
    return dict(f=f.reshape(n, 3), fs=fs.reshape(n, 3), e_vdw=en[0], e_elec=en[1], e_slow=en[2],
                virial=vir.reshape(3, 3), virial_slow=virs.reshape(3, 3), npairs=int(npairs[0]))


ALCH_KEYS = ("elecLambdaUp", "elecLambdaDown", "vdwLambdaUp", "vdwLambdaDown", "elecLambda2Up", "elecLambda2Down",
             "vdwLambda2Up", "vdwLambda2Down", "vdwShiftUp", "vdwShiftDown", "vdwShift2Up", "vdwShift2Down")


def nb_fp64_alch(pos, q_scaled, vdw_type, A, B, excl_ptr, excl_idx, box, cutoff, switchdist, beta, partition,
                 alch, nthreads=0, vdw_force_switching=False):
    """FP64 alchemical oracle.  alch: mapping with ALCH_KEYS, 'alchDecouple'
    and optionally 'alchVdwShiftCoeff' (an AlchData works); vdw_force_switching
    selects NAMD's force-switched soft-core LJ.  Returns nb_fp64's
    dict plus the lambda2 energies e_vdw_s, e_elec_s, e_slow_s and the TI
    derivatives e_vdw_ti_1/2, e_elec_ti_1/2, e_slow_ti_1/2 (1: partition 1,
    2: partition 2)."""
    n = len(q_scaled)
    ntypes = int(round(np.sqrt(np.asarray(A).size)))
    get = (lambda k: alch[k]) if isinstance(alch, dict) else (lambda k: getattr(alch, k))
    has = (lambda k: k in alch) if isinstance(alch, dict) else (lambda k: hasattr(alch, k))
    lam = np.array([float(get(k)) for k in ALCH_KEYS] +
                   [float(get("alchVdwShiftCoeff")) if has("alchVdwShiftCoeff") else 0.0], np.float64)
    f = np.zeros(3 * n)
    fs = np.zeros(3 * n)
    en = np.zeros(12)
    vir = np.zeros(9)
    virs = np.zeros(9)
    npairs = np.zeros(1, np.int64)
    fn = lib().oracle_nb_fp64_alch_vdw_force_switch if vdw_force_switching else lib().oracle_nb_fp64_alch
    fn(n, _c(pos, np.float64).reshape(-1), _c(q_scaled, np.float64), _c(vdw_type, np.int32),
                              ntypes, _c(A, np.float64).reshape(-1), _c(B, np.float64).reshape(-1),
                              _c(excl_ptr, np.int64), _c(excl_idx, np.int32), _c(box, np.float64), cutoff,
                              switchdist, beta, _c(partition, np.int8), lam, int(bool(get("alchDecouple"))), f, fs,
                              en, vir, virs, npairs, int(nthreads))
    return dict(f=f.reshape(n, 3), fs=fs.reshape(n, 3), e_vdw=en[0], e_elec=en[1], e_slow=en[2], e_vdw_s=en[3],
                e_elec_s=en[4], e_slow_s=en[5], e_vdw_ti_1=en[6], e_vdw_ti_2=en[7], e_elec_ti_1=en[8],
                e_elec_ti_2=en[9], e_slow_ti_1=en[10], e_slow_ti_2=en[11], virial=vir.reshape(3, 3),
                virial_slow=virs.reshape(3, 3), npairs=int(npairs[0]))


def nb_fp64_cell(pos, q_scaled, vdw_type, A, B, excl_ptr, excl_idx, cell, cutoff, switchdist, beta, nthreads=0,
                 vdw_force_switching=False):
    """nb_fp64 in a general periodic cell (cell: 3 x 3 lattice rows a, b, c),
    by brute force over all pairs and the 27 images around the minimum image."""
    n = len(q_scaled)
    ntypes = int(round(np.sqrt(np.asarray(A).size)))
    f = np.zeros(3 * n)
    fs = np.zeros(3 * n)
    en = np.zeros(3)
    vir = np.zeros(9)
    virs = np.zeros(9)
    npairs = np.zeros(1, np.int64)
    lib().oracle_nb_fp64_cell(n, _c(pos, np.float64).reshape(-1), _c(q_scaled, np.float64), _c(vdw_type, np.int32),
                              ntypes, _c(A, np.float64).reshape(-1), _c(B, np.float64).reshape(-1),
                              _c(excl_ptr, np.int64), _c(excl_idx, np.int32), _c(cell, np.float64).reshape(-1),
                              cutoff, switchdist, beta, int(bool(vdw_force_switching)), f, fs, en, vir, virs,
                              npairs, int(nthreads))
    return dict(f=f.reshape(n, 3), fs=fs.reshape(n, 3), e_vdw=en[0], e_elec=en[1], e_slow=en[2],
                virial=vir.reshape(3, 3), virial_slow=virs.reshape(3, 3), npairs=int(npairs[0]))


def mod_excl_fp64(pos, q_scaled, vdw_type, A, B, A14, B14, mod_ptr, mod_idx, box, cutoff, switchdist, scale14,
                  vdw_force_switching=False, partition=None, alch=None, cell=None):
    """FP64 modified (scaled 1-4) exclusion correction over the listed pairs
    (atom CSR mod_ptr / mod_idx): returns f, fs (n x 3), e_vdw, e_elec, e_slow,
    virial, virial_slow, npairs (pairs within the cutoff).  With partition and
    alch (as nb_fp64_alch takes them) the alchemical correction, and also the
    lambda2 sums e_*_s and the TI sums e_*_ti_1/2.  cell (3 x 3 lattice rows):
    a general cell instead of the orthorhombic box."""
    n = len(q_scaled)
    ntypes = int(round(np.sqrt(np.asarray(A).size)))
    f = np.zeros(3 * n)
    fs = np.zeros(3 * n)
    en = np.zeros(12)
    vir = np.zeros(9)
    virs = np.zeros(9)
    npairs = np.zeros(1, np.int64)
    args = (n, _c(pos, np.float64).reshape(-1), _c(q_scaled, np.float64), _c(vdw_type, np.int32), ntypes,
            _c(A, np.float64).reshape(-1), _c(B, np.float64).reshape(-1), _c(A14, np.float64).reshape(-1),
            _c(B14, np.float64).reshape(-1), _c(mod_ptr, np.int64), _c(mod_idx, np.int32),
            _c(box if box is not None else np.zeros(3), np.float64),
            cutoff, switchdist, scale14)
    out = dict()
    if cell is not None:
        if partition is not None:
            raise ValueError("the general-cell modified-exclusion oracle has no alchemical route")
        args = args[:11] + (_c(cell, np.float64).reshape(-1),) + args[12:]
        lib().oracle_mod_excl_cell_fp64(*args, int(bool(vdw_force_switching)), f, fs, en, vir, virs, npairs)
    elif partition is None:
        lib().oracle_mod_excl_fp64(*args, int(bool(vdw_force_switching)), f, fs, en, vir, virs, npairs)
    else:
        if vdw_force_switching:
            raise ValueError("the alchemical modified-exclusion correction uses potential switching only")
        get = (lambda k: alch[k]) if isinstance(alch, dict) else (lambda k: getattr(alch, k))
        has = (lambda k: k in alch) if isinstance(alch, dict) else (lambda k: hasattr(alch, k))
        lam = np.array([float(get(k)) for k in ALCH_KEYS] +
                       [float(get("alchVdwShiftCoeff")) if has("alchVdwShiftCoeff") else 0.0], np.float64)
        lib().oracle_mod_excl_alch_fp64(*args, _c(partition, np.int8), lam, int(bool(get("alchDecouple"))), f, fs,
                                        en, vir, virs, npairs)
        out = dict(e_vdw_s=en[3], e_elec_s=en[4], e_slow_s=en[5], e_vdw_ti_1=en[6], e_vdw_ti_2=en[7],
                   e_elec_ti_1=en[8], e_elec_ti_2=en[9], e_slow_ti_1=en[10], e_slow_ti_2=en[11])
    return dict(f=f.reshape(n, 3), fs=fs.reshape(n, 3), e_vdw=en[0], e_elec=en[1], e_slow=en[2],
                virial=vir.reshape(3, 3), virial_slow=virs.reshape(3, 3), npairs=int(npairs[0]), **out)


def nb_adjoint_fp64(pos, q_scaled, vdw_type, ntypes, excl_ptr, excl_idx, box, cutoff, switchdist, beta, g,
                    nthreads=0):
    """FP64 force-matching adjoint: given g = dL/dF (n x 3, total force),
    returns (dL/dA, dL/dB as ntypes x ntypes ordered bins, dL/dq_scaled)."""
    n = len(q_scaled)
    gA = np.zeros(ntypes * ntypes)
    gB = np.zeros(ntypes * ntypes)
    gq = np.zeros(n)
    lib().oracle_nb_adjoint_fp64(n, _c(pos, np.float64).reshape(-1), _c(q_scaled, np.float64),
                                 _c(vdw_type, np.int32), int(ntypes), _c(excl_ptr, np.int64), _c(excl_idx, np.int32),
                                 _c(box, np.float64), cutoff, switchdist, beta, _c(g, np.float64).reshape(-1), gA, gB,
                                 gq, int(nthreads))
    return gA.reshape(ntypes, ntypes), gB.reshape(ntypes, ntypes), gq


def pairs_fp32(pos, excl_ptr, excl_idx, box, cutoff2f, plcutoff2f, want_pairs=False, maxpairs=None):
    """Canonical FP32 pair sets (see nb_oracle.c). Returns dict of counts, checksums, optional pairs."""
    n = pos.shape[0]
    out = np.zeros(4, np.int64)
    sums = np.zeros(2, np.uint64)
    xors = np.zeros(2, np.uint64)
    buf = None
    cap = 0
    if want_pairs:
        cap = int(maxpairs) if maxpairs else 200 * n
        buf = np.zeros((cap, 2), np.int32)
    lib().oracle_pairs_fp32(n, _c(pos, np.float64).reshape(-1), _c(excl_ptr, np.int64), _c(excl_idx, np.int32),
                            _c(box, np.float64), np.float32(cutoff2f), np.float32(plcutoff2f), out, sums, xors,
                            buf.ctypes.data if buf is not None else None, cap)
    res = dict(n_cut=int(out[0]), n_list=int(out[1]), n_excl_list=int(out[2]),
               sum_cut=int(sums[0]), xor_cut=int(xors[0]), sum_list=int(sums[1]), xor_list=int(xors[1]))
    if want_pairs:
        if out[3] < out[1]:
            raise ValueError("pair buffer too small")
        res["pairs"] = buf[: out[3]].copy()
    return res


def pairs_exact(pos, excl_ptr, excl_idx, box, cutoff, pairlistdist, want_pairs=False, maxpairs=None):
    """NAMD CPU membership rule (r2 <= rc^2 in FP64, exact on effective positions); same outputs as pairs_fp32."""
    n = pos.shape[0]
    out = np.zeros(4, np.int64)
    sums = np.zeros(2, np.uint64)
    xors = np.zeros(2, np.uint64)
    buf = None
    cap = 0
    if want_pairs:
        cap = int(maxpairs) if maxpairs else 200 * n
        buf = np.zeros((cap, 2), np.int32)
    lib().oracle_pairs_exact(n, _c(pos, np.float64).reshape(-1), _c(excl_ptr, np.int64), _c(excl_idx, np.int32),
                             _c(box, np.float64), float(cutoff) ** 2, float(pairlistdist) ** 2, out, sums, xors,
                             buf.ctypes.data if buf is not None else None, cap)
    res = dict(n_cut=int(out[0]), n_list=int(out[1]), n_excl_list=int(out[2]),
               sum_cut=int(sums[0]), xor_cut=int(xors[0]), sum_list=int(sums[1]), xor_list=int(xors[1]))
    if want_pairs:
        if out[3] < out[1]:
            raise ValueError("pair buffer too small")
        res["pairs"] = buf[: out[3]].copy()
    return res


def namd_tables(cutoff, switchdist, ewaldcof, limitdist=0.0):
    """NAMD's FP64 spline tables: (r2_table[n], tables[5, n, 4]) for fast, scor, slow, vdwa, vdwb."""
    n = C.c_int(0)
    lib().oracle_namd_tables(cutoff, switchdist, ewaldcof, limitdist, C.byref(n), None, None)
    r2 = np.zeros(n.value)
    t = np.zeros(5 * 4 * n.value)
    lib().oracle_namd_tables(cutoff, switchdist, ewaldcof, limitdist, C.byref(n), r2.ctypes.data, t.ctypes.data)
    return r2, t.reshape(5, n.value, 4)


def nb_table(pos, q_scaled, vdw_type, A, B, excl_ptr, excl_idx, box, cutoff, switchdist, force4, energy4,
             nthreads=0):
    """Table-mode pair evaluation with the given (N, 4) float32 tables; outputs as nb_fp64."""
    n = len(q_scaled)
    ntypes = int(round(np.sqrt(np.asarray(A).size)))
    ft = _c(force4, np.float32).reshape(-1)
    et = _c(energy4, np.float32).reshape(-1)
    f = np.zeros(3 * n)
    fs = np.zeros(3 * n)
    en = np.zeros(3)
    vir = np.zeros(9)
    virs = np.zeros(9)
    npairs = np.zeros(1, np.int64)
    lib().oracle_nb_table(n, _c(pos, np.float64).reshape(-1), _c(q_scaled, np.float64), _c(vdw_type, np.int32),
                          ntypes, _c(A, np.float64).reshape(-1), _c(B, np.float64).reshape(-1),
                          _c(excl_ptr, np.int64), _c(excl_idx, np.int32), _c(box, np.float64), cutoff, switchdist,
                          ft.size // 4, ft, et, f, fs, en, vir, virs, npairs, int(nthreads))
    return dict(f=f.reshape(n, 3), fs=fs.reshape(n, 3), e_vdw=en[0], e_elec=en[1], e_slow=en[2],
                virial=vir.reshape(3, 3), virial_slow=virs.reshape(3, 3), npairs=int(npairs[0]))


def namd_gpu_tables(n, cutoff, switchdist, ewaldcof):
    """NAMD's float4 1/r force and energy tables (n entries each) from the spline restatement."""
    f = np.zeros(4 * n, np.float32)
    e = np.zeros(4 * n, np.float32)
    lib().oracle_namd_gpu_tables(int(n), cutoff, switchdist, ewaldcof, f, e)
    return f.reshape(n, 4), e.reshape(n, 4)


def namd_cpu(pos, q_scaled, vdw_type, A, B, excl_ptr, excl_idx, box, cutoff, switchdist, beta, nthreads=0):
    """NAMD CPU nonbonded path restatement (spline tables); same outputs as nb_fp64."""
    n = len(q_scaled)
    ntypes = int(round(np.sqrt(np.asarray(A).size)))
    f = np.zeros(3 * n)
    fs = np.zeros(3 * n)
    en = np.zeros(3)
    vir = np.zeros(9)
    virs = np.zeros(9)
    npairs = np.zeros(1, np.int64)
    lib().oracle_namd_cpu(n, _c(pos, np.float64).reshape(-1), _c(q_scaled, np.float64), _c(vdw_type, np.int32),
                          ntypes, _c(A, np.float64).reshape(-1), _c(B, np.float64).reshape(-1),
                          _c(excl_ptr, np.int64), _c(excl_idx, np.int32), _c(box, np.float64), cutoff, switchdist,
                          beta, f, fs, en, vir, virs, npairs, int(nthreads))
    return dict(f=f.reshape(n, 3), fs=fs.reshape(n, 3), e_vdw=en[0], e_elec=en[1], e_slow=en[2],
                virial=vir.reshape(3, 3), virial_slow=virs.reshape(3, 3), npairs=int(npairs[0]))


def pair_hash(i, j, n):
    """Vectorised oracle_pair_hash (splitmix64 of min*n+max) in numpy uint64 arithmetic."""
    i = np.asarray(i, np.int64)
    j = np.asarray(j, np.int64)
    a = np.minimum(i, j).astype(np.uint64)
    b = np.maximum(i, j).astype(np.uint64)
    with np.errstate(over="ignore"):
        x = a * np.uint64(n) + b
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x


def pme_spread_gpu(atoms_f4, K, order):
    K1, K2, K3 = K
    g = np.zeros(K1 * K2 * K3)
    lib().oracle_pme_spread_gpu(atoms_f4.shape[0], _c(atoms_f4, np.float32).reshape(-1), K1, K2, K3, order, g)
    return g.reshape(K3, K2, K1)


def pme_gather_gpu(atoms_f4, K, order, grid, recip=None):
    K1, K2, K3 = K
    n = atoms_f4.shape[0]
    f = np.zeros(3 * n)
    r = None if recip is None else _c(recip, np.float64).reshape(-1)
    lib().oracle_pme_gather_gpu(n, _c(atoms_f4, np.float32).reshape(-1), K1, K2, K3, order,
                                _c(grid, np.float64).reshape(-1), r.ctypes.data if r is not None else None, f)
    return f.reshape(n, 3)


def pme_scale_coordinates(pos, origin, recip, K, order):
    n = pos.shape[0]
    p = np.zeros(3 * n)
    lib().oracle_pme_scale_coordinates(n, _c(pos, np.float64).reshape(-1), _c(origin, np.float64),
                                       _c(recip, np.float64).reshape(-1), K[0], K[1], K[2], order, p)
    return p.reshape(n, 3)


def pme_spread_cpu(p, q, K, order):
    K1, K2, K3 = K
    g = np.zeros(K1 * K2 * K3)
    lib().oracle_pme_spread_cpu(p.shape[0], _c(p, np.float64).reshape(-1), _c(q, np.float64), K1, K2, K3, order, g)
    return g.reshape(K3, K2, K1)


def pme_gather_cpu(p, q, K, order, grid):
    K1, K2, K3 = K
    n = p.shape[0]
    f = np.zeros(3 * n)
    lib().oracle_pme_gather_cpu(n, _c(p, np.float64).reshape(-1), _c(q, np.float64), K1, K2, K3, order,
                                _c(grid, np.float64).reshape(-1), f)
    return f.reshape(n, 3)


def bspline_gpu(w, order):
    th = np.zeros(order)
    dth = np.zeros(order)
    lib().oracle_bspline_gpu(float(w), order, th, dth)
    return th, dth


# ---- PME reciprocal space (FP64, numpy's FFT) --------------------------------

def pme_spline_at_integers(order):
    """NAMD's compute_b_spline at fraction 0 (src/PmeBase.inl:86-125): the
    order-n cardinal B-spline at u = 1 .. n, by its own recursion (start at
    n = 3, raise the order one step at a time)."""
    x = 0.0
    M = np.zeros(order + 1)          # 1-based like the reference
    M[1] = 0.5 * (1.0 - x) ** 2
    M[2] = (1.0 - x) * x + 0.5
    M[3] = 0.5 * x * x
    M[order] = 0.0 if order > 3 else M[order]
    for n in range(4, order + 1):
        div = 1.0 / (n - 1)
        M[n] = x * div * M[n - 1]
        for j in range(1, n - 1):
            M[n - j] = ((x + j) * M[n - j - 1] + (n - x - j) * M[n - j]) * div
        M[1] *= (1.0 - x) * div
    return M[1:order + 1].copy()


def pme_b_moduli(K, order):
    """compute_b_moduli (src/PmeKSpace.C:21-61): bm[k] = 1/|DFT(M)[k]|^2 with
    dftmod's cos/sin sums."""
    M = np.zeros(K)
    M[:order] = pme_spline_at_integers(order)
    j = np.arange(K)
    mod = np.empty(K)
    for k in range(K):
        arg = 2.0 * np.pi * k * j / K
        mod[k] = (M * np.cos(arg)).sum() ** 2 + (M * np.sin(arg)).sum() ** 2
    return 1.0 / mod


def _kspace(K, recip, volume, beta, order):
    """Per half-complex point (k3, k2, k1): theta = bm1 bm2 bm3 exp(-pi^2 m^2/beta^2)/(pi V m^2),
    m (3 components), m^2 and the multiplicity (2, or 1 on the k1 = 0 and K1/2 planes)."""
    K1, K2, K3 = K
    K1h = K1 // 2 + 1
    recip = np.asarray(recip, np.float64).reshape(3, 3)
    k1 = np.arange(K1h)
    k2 = np.arange(K2)
    k3 = np.arange(K3)
    k2s = np.where(k2 <= K2 // 2, k2, k2 - K2)
    k3s = np.where(k3 <= K3 // 2, k3, k3 - K3)
    m = (k1[None, None, :, None] * recip[0] + k2s[None, :, None, None] * recip[1]
         + k3s[:, None, None, None] * recip[2])
    msq = (m * m).sum(axis=-1)
    msq[0, 0, 0] = 1.0
    b1, b2, b3 = pme_b_moduli(K1, order), pme_b_moduli(K2, order), pme_b_moduli(K3, order)
    bm = b3[:, None, None] * b2[None, :, None] * b1[None, None, :K1h]
    fac = np.pi ** 2 / beta ** 2
    theta = bm * np.exp(-fac * msq) / (np.pi * volume * msq)
    theta[0, 0, 0] = 0.0
    mult = np.full(K1h, 2.0)
    mult[0] = 1.0
    if K1 % 2 == 0:
        mult[K1 // 2] = 1.0
    return theta, m, msq, mult[None, None, :], fac


def pme_scalar_sum(trans, K, recip, volume, beta, order):
    """scalar_sum (src/CudaPmeSolverUtilKernel.cu:73-320) in FP64 on a
    half-complex grid (K3, K2, K1/2+1): returns (multiplied grid, energy,
    virial6 {xx, xy, xz, yy, yz, zz}), energy = 1/2 sum mult*theta*|S|^2,
    virial = -1/2 sum mult*theta*|S|^2 (2 (fac + 1/m^2) m_a m_b - delta_ab)."""
    theta, m, msq, mult, fac = _kspace(K, recip, volume, beta, order)
    S2 = np.abs(trans) ** 2
    f = mult * theta * S2
    energy = 0.5 * f.sum()
    vir = 2.0 * (fac + 1.0 / msq)
    v = []
    for a, b in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)):
        t = vir * m[..., a] * m[..., b] - (1.0 if a == b else 0.0)
        v.append(-0.5 * (f * t).sum())
    return trans * theta, energy, np.array(v)


def pme_recip(grid, K, recip, volume, beta, order):
    """FP64 reciprocal solve of CudaPmeOneDevice::compute
    (src/CudaPmeSolverUtil.C:1579-1620): R2C FFT of grid (K3, K2, K1),
    scalar_sum, unnormalised C2R.  Returns (potential grid, energy, virial6)."""
    K1, K2, K3 = K
    trans = np.fft.rfftn(np.asarray(grid, np.float64).reshape(K3, K2, K1))
    out, energy, vir = pme_scalar_sum(trans, K, recip, volume, beta, order)
    phi = np.fft.irfftn(out, s=(K3, K2, K1), axes=(0, 1, 2)) * (K1 * K2 * K3)
    return phi, energy, vir


def ewald_recip_direct(pos, q_scaled, cell, beta, kmax):
    """Brute-force Ewald reciprocal sum (no grid): E = 1/(2 pi V) sum_{m != 0}
    exp(-pi^2 m^2/beta^2)/m^2 |S(m)|^2, S(m) = sum_j q_j exp(2 pi i m.r_j),
    m = n1 a* + n2 b* + n3 c*, |n_i| <= kmax; and F_j = -dE/dr_j."""
    cell = np.asarray(cell, np.float64).reshape(3, 3)
    recip = np.linalg.inv(cell).T
    V = abs(np.linalg.det(cell))
    pos = np.asarray(pos, np.float64)
    q = np.asarray(q_scaled, np.float64)
    rng = np.arange(-kmax, kmax + 1)
    n = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    n = n[np.any(n != 0, axis=1)]
    m = n @ recip
    msq = (m * m).sum(1)
    f = np.exp(-np.pi ** 2 * msq / beta ** 2) / msq
    E = 0.0
    F = np.zeros_like(pos)
    for c0 in range(0, m.shape[0], 4096):
        mm = m[c0:c0 + 4096]
        ph = 2.0 * np.pi * (pos @ mm.T)
        e = np.exp(1j * ph)
        S = (q[:, None] * e).sum(0)
        ff = f[c0:c0 + 4096]
        E += (ff * np.abs(S) ** 2).sum()
        # dE/dr_j = 1/(2 pi V) sum f * 2 Re(conj(S) q_j 2 pi i m e_j)
        g = (ff[None, :] * np.real(np.conj(S)[None, :] * 1j * e)) @ mm
        F -= (2.0 / V) * q[:, None] * g
    return E / (2.0 * np.pi * V), F


def pme_exclusion_correction(pos, q_scaled, excl_ptr, excl_idx, box, beta, cutoff):
    """PME exclusion correction (exclusionForce, src/ComputeBondedCUDAKernel.cu:633-772)
    in FP64: for each fully excluded pair (once) with minimum-image r^2 <
    cutoff^2, E = -qq erf(beta r)/r.  Returns (E, F (n,3), virial 3x3 =
    sum f_i (x) r_ij)."""
    from scipy.special import erf
    pos = np.asarray(pos, np.float64)
    q = np.asarray(q_scaled, np.float64)
    box = np.asarray(box, np.float64)
    n = len(q)
    ii = np.repeat(np.arange(n), np.diff(np.asarray(excl_ptr)))
    jj = np.asarray(excl_idx, np.int64)
    a, b = np.minimum(ii, jj), np.maximum(ii, jj)
    key = np.unique(a * n + b)
    a, b = key // n, key % n
    d = pos[a] - pos[b]
    d -= box * np.rint(d / box)
    r2 = (d * d).sum(1)
    keep = r2 < cutoff * cutoff
    a, b, d, r2 = a[keep], b[keep], d[keep], r2[keep]
    r = np.sqrt(r2)
    qq = q[a] * q[b]
    e = erf(beta * r) / r
    E = float(-(qq * e).sum())
    dedr = -qq * (2.0 / np.sqrt(np.pi) * beta * np.exp(-beta * beta * r2) / r - e / r)
    f = (-dedr / r)[:, None] * d
    F = np.zeros((n, 3))
    np.add.at(F, a, f)
    np.add.at(F, b, -f)
    W = np.einsum("ka,kb->ab", f, d)
    return E, F, W


# ---- rigid water constraints (FP64) -------------------------------------------

def settle_geometry(mO, mH, rOH, rHH):
    """Canonical triangle in the centre-of-mass frame (settle1init,
    src/Settle.C:46-61): O at (0, ra), H at (-+rc, -rb)."""
    rc = 0.5 * rHH
    h = np.sqrt(rOH * rOH - rc * rc)
    ra = h * 2.0 * mH / (mO + 2.0 * mH)
    return ra, h - ra, rc


def settle_positions(ref, pos, mO, mH, rOH, rHH):
    """SETTLE (settle1, src/Settle.C:63-290; Miyamoto & Kollman 1992) for
    waters given as (nw, 3, 3) arrays [water][O, H1, H2][xyz]: returns the
    constrained positions."""
    ra, rb, rc = settle_geometry(mO, mH, rOH, rHH)
    M = mO + 2 * mH
    A0, B0, C0 = ref[:, 0], ref[:, 1], ref[:, 2]
    A1, B1, C1 = pos[:, 0], pos[:, 1], pos[:, 2]
    d0 = (mO * A1 + mH * (B1 + C1)) / M
    a1, b1, c1 = A1 - d0, B1 - d0, C1 - d0
    b0, c0 = B0 - A0, C0 - A0

    def unit(v):
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    ez = unit(np.cross(b0, c0))
    ex = unit(np.cross(a1, ez))
    ey = np.cross(ez, ex)

    def dot(u, v):
        return (u * v).sum(1)

    xb0, yb0, xc0, yc0 = dot(b0, ex), dot(b0, ey), dot(c0, ex), dot(c0, ey)
    za1 = dot(a1, ez)
    xb1, yb1, zb1 = dot(b1, ex), dot(b1, ey), dot(b1, ez)
    xc1, yc1, zc1 = dot(c1, ex), dot(c1, ey), dot(c1, ez)
    sphi = za1 / ra
    cphi = np.sqrt(np.maximum(0.0, 1 - sphi ** 2))
    spsi = (zb1 - zc1) / (2 * rc * cphi)
    cpsi = np.sqrt(np.maximum(0.0, 1 - spsi ** 2))
    ya2 = ra * cphi
    xb2 = -rc * cpsi
    yb2 = -rb * cphi - rc * spsi * sphi
    yc2 = -rb * cphi + rc * spsi * sphi
    al = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2
    be = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2
    ga = xb0 * yb1 - xb1 * yb0 + xc0 * yc1 - xc1 * yc0
    ab2 = al * al + be * be
    sth = (al * ga - be * np.sqrt(np.maximum(0.0, ab2 - ga * ga))) / ab2
    cth = np.sqrt(np.maximum(0.0, 1 - sth ** 2))
    out = np.empty_like(pos)
    comps = [(-ya2 * sth, ya2 * cth, ra * sphi),
             (xb2 * cth - yb2 * sth, xb2 * sth + yb2 * cth, -rb * sphi + rc * spsi * cphi),
             (-xb2 * cth - yc2 * sth, -xb2 * sth + yc2 * cth, -rb * sphi - rc * spsi * cphi)]
    for k, (x, y, z) in enumerate(comps):
        out[:, k] = d0 + x[:, None] * ex + y[:, None] * ey + z[:, None] * ez
    return out


def shake_positions(ref, pos, mO, mH, rOH, rHH, tol=1e-14, iters=10000):
    """Iterative SHAKE (Ryckaert et al. 1977) for the same rigid waters:
    corrections along the reference bond vectors, mass weighted, until every
    distance matches to tol.  An independent solver of the problem SETTLE
    solves in closed form."""
    x = pos.copy()
    m = np.array([mO, mH, mH])
    bonds = [(0, 1, rOH), (0, 2, rOH), (1, 2, rHH)]
    for _ in range(iters):
        worst = 0.0
        for i, j, d in bonds:
            rij = x[:, i] - x[:, j]
            r0 = ref[:, i] - ref[:, j]
            diff = (rij * rij).sum(1) - d * d
            worst = max(worst, np.abs(diff).max())
            g = diff / (2.0 * (1 / m[i] + 1 / m[j]) * (rij * r0).sum(1))
            x[:, i] -= (g / m[i])[:, None] * r0
            x[:, j] += (g / m[j])[:, None] * r0
        if worst < tol:
            break
    return x


def settle_velocities(pos, vel, mO, mH):
    """settle2 (src/Settle.C:1420-1486) restated: bond impulses making the
    three distances of each water stationary, from the 3x3 system solved by
    numpy.  pos, vel: (nw, 3, 3)."""
    m = np.array([mO, mH, mH])
    bonds = [(0, 1), (0, 2), (1, 2)]
    out = vel.copy()
    for w in range(pos.shape[0]):
        r = [pos[w, i] - pos[w, j] for i, j in bonds]
        Mx = np.zeros((3, 3))
        b = np.zeros(3)
        for k, (i, j) in enumerate(bonds):
            b[k] = -np.dot(vel[w, i] - vel[w, j], r[k])
            for l, (p, q) in enumerate(bonds):
                # d(v_i - v_j)/d lam_l . r_k
                ci = (1 / m[i] if p == i else 0) - (1 / m[i] if q == i else 0)
                cj = (1 / m[j] if p == j else 0) - (1 / m[j] if q == j else 0)
                Mx[k, l] = (ci - cj) * np.dot(r[l], r[k])
        lam = np.linalg.solve(Mx, b)
        for l, (p, q) in enumerate(bonds):
            out[w, p] += lam[l] * r[l] / m[p]
            out[w, q] -= lam[l] * r[l] / m[q]
    return out


# ---- harmonic bonds and angles (FP64) -----------------------------------------

def bonded_fp64(pos, box, bonds, bond_k, bond_r0, angles, angle_k, angle_theta0):
    """bondForce / angleForce (src/ComputeBondedCUDAKernel.cu:194-280,
    :775-900) in CHARMM form, FP64: returns (E_bond, E_angle, F (n, 3))."""
    pos = np.asarray(pos, np.float64)
    box = np.asarray(box, np.float64)
    F = np.zeros_like(pos)
    d = pos[bonds[:, 0]] - pos[bonds[:, 1]]
    d -= box * np.rint(d / box)
    r = np.linalg.norm(d, axis=1)
    db = r - bond_r0
    Eb = float((bond_k * db * db).sum())
    g = (-2.0 * bond_k * db / r)[:, None] * d
    np.add.at(F, bonds[:, 0], g)
    np.add.at(F, bonds[:, 1], -g)
    u = pos[angles[:, 0]] - pos[angles[:, 1]]
    v = pos[angles[:, 2]] - pos[angles[:, 1]]
    u -= box * np.rint(u / box)
    v -= box * np.rint(v / box)
    ru = np.linalg.norm(u, axis=1)
    rv = np.linalg.norm(v, axis=1)
    cs = np.clip((u * v).sum(1) / (ru * rv), -1.0, 1.0)
    th = np.arccos(cs)
    sn = np.maximum(np.sqrt(1.0 - cs * cs), 1e-8)
    dE = 2.0 * angle_k * (th - angle_theta0)
    fi = (dE / sn)[:, None] * (v / (ru * rv)[:, None] - (cs / (ru * ru))[:, None] * u)
    fk = (dE / sn)[:, None] * (u / (ru * rv)[:, None] - (cs / (rv * rv))[:, None] * v)
    np.add.at(F, angles[:, 0], fi)
    np.add.at(F, angles[:, 2], fk)
    np.add.at(F, angles[:, 1], -(fi + fk))
    Ea = float((angle_k * (th - angle_theta0) ** 2).sum())
    return Eb, Ea, F
