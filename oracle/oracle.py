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
    return _LIB


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def nb_fp64(pos, q_scaled, vdw_type, A, B, excl_ptr, excl_idx, box, cutoff, switchdist, beta, nthreads=0):
    """FP64 analytic oracle. Returns dict(f, fs, e_vdw, e_elec, e_slow, virial, virial_slow, npairs)."""
    n = len(q_scaled)
    ntypes = int(round(np.sqrt(np.asarray(A).size)))
    f = np.zeros(3 * n)
    fs = np.zeros(3 * n)
    en = np.zeros(3)
    vir = np.zeros(9)
    virs = np.zeros(9)
    npairs = np.zeros(1, np.int64)
    lib().oracle_nb_fp64(n, _c(pos, np.float64).reshape(-1), _c(q_scaled, np.float64), _c(vdw_type, np.int32),
                         ntypes, _c(A, np.float64).reshape(-1), _c(B, np.float64).reshape(-1),
                         _c(excl_ptr, np.int64), _c(excl_idx, np.int32), _c(box, np.float64), cutoff, switchdist,
                         beta, f, fs, en, vir, virs, npairs, int(nthreads))
    return dict(f=f.reshape(n, 3), fs=fs.reshape(n, 3), e_vdw=en[0], e_elec=en[1], e_slow=en[2],
                virial=vir.reshape(3, 3), virial_slow=virs.reshape(3, 3), npairs=int(npairs[0]))


ALCH_KEYS = ("elecLambdaUp", "elecLambdaDown", "vdwLambdaUp", "vdwLambdaDown", "elecLambda2Up", "elecLambda2Down",
             "vdwLambda2Up", "vdwLambda2Down", "vdwShiftUp", "vdwShiftDown", "vdwShift2Up", "vdwShift2Down")


def nb_fp64_alch(pos, q_scaled, vdw_type, A, B, excl_ptr, excl_idx, box, cutoff, switchdist, beta, partition,
                 alch, nthreads=0):
    """FP64 alchemical (FEP) oracle.  alch: mapping with ALCH_KEYS and
    'alchDecouple' (an AlchData works).  Returns nb_fp64's dict plus the
    lambda2 energies e_vdw_s, e_elec_s, e_slow_s."""
    n = len(q_scaled)
    ntypes = int(round(np.sqrt(np.asarray(A).size)))
    get = (lambda k: alch[k]) if isinstance(alch, dict) else (lambda k: getattr(alch, k))
    lam = np.array([float(get(k)) for k in ALCH_KEYS], np.float64)
    f = np.zeros(3 * n)
    fs = np.zeros(3 * n)
    en = np.zeros(6)
    vir = np.zeros(9)
    virs = np.zeros(9)
    npairs = np.zeros(1, np.int64)
    lib().oracle_nb_fp64_alch(n, _c(pos, np.float64).reshape(-1), _c(q_scaled, np.float64), _c(vdw_type, np.int32),
                              ntypes, _c(A, np.float64).reshape(-1), _c(B, np.float64).reshape(-1),
                              _c(excl_ptr, np.int64), _c(excl_idx, np.int32), _c(box, np.float64), cutoff,
                              switchdist, beta, _c(partition, np.int8), lam, int(bool(get("alchDecouple"))), f, fs,
                              en, vir, virs, npairs, int(nthreads))
    return dict(f=f.reshape(n, 3), fs=fs.reshape(n, 3), e_vdw=en[0], e_elec=en[1], e_slow=en[2], e_vdw_s=en[3],
                e_elec_s=en[4], e_slow_s=en[5], virial=vir.reshape(3, 3), virial_slow=virs.reshape(3, 3),
                npairs=int(npairs[0]))


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
