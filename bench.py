#!/usr/bin/env python
"""Fock-build throughput of the MI355X engine on the seeded (H2O)_8 6-31G**
cluster (200 Cartesian functions, Schwarz cutoff 1e-12, FP64).

A step is one direct Fock build (J and K from one density, pq_fock_jk) over
the full screened quartet list, replayed from a HIP graph capture of that call
(--eager times plain calls).  The class kernels run concurrently on the
engine's streams; their bra-run lengths come from the stored autotuner table
(pyquante_amd/tuning.json, --tune).  Prints one JSON line.

The roofline block is for the dominant class kernel: its algorithmic Rys
flops per launch (the F_prim/F_dig operation count over the class's
primitive and contracted quartets) over its average launch duration, timed
with HIP events on the stream the kernel is launched on, in the same
concurrent configuration as the timed builds, against the MI355X FP64 vector
peak.  `traffic` is the HBM bytes per launch of that kernel from rocprofv3
FETCH_SIZE/WRITE_SIZE runs (profiles/traffic.json), when recorded;
`counted_fp64_tflops` is its FP64 VALU work counted by rocprofv3 (64 lanes x
(2 FMA + ADD + MUL) per launch, profiles/valu.json) over the same launch
duration.  The cpu_baseline block times the compiled reference crys.c (oracle/_ref) on a
seeded sample of the same workload's unique function quartets,
single-threaded; cpu_baseline_all_cores runs the same sample in one process
per host core (up to 16) and sums their rates.

Outside the timed region the line also carries: (H2O)_16 FP64 against mixed
precision (times, energy error), forces per second on (H2O)_8, J/K of two
densities in one pass against two builds (the unrestricted path), MP2 on
(H2O)_8 (stored ERIs, AO->MO transform as device GEMMs), the batched
workload (512 seeded (H2O)_n, n in 1..12, in one engine: Fock builds, SCF and
forces in FP64 and mixed precision, the largest per-molecule errors), and the HGP
path on (H2O)_8 (all 202,015,050 unique ERIs by the shell-quartet GPU HGP
kernel and by the per-function one, against the Rys kernels, RHF energies
through both, the compiled reference chgp.c on a CPU sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6


def _basis_arrays(bfs):
    """Flat primitive arrays of a basis in the reference batch harness's layout."""
    import numpy as np
    nprim = np.array([len(f.pexps) for f in bfs], dtype=np.int32)
    return {"nprim": nprim, "poff": np.concatenate([[0], np.cumsum(nprim)[:-1]]).astype(np.int32),
            "e": np.array([x for f in bfs for x in f.pexps]), "c": np.array([x for f in bfs for x in f.pcoefs]),
            "nr": np.array([x for f in bfs for x in f.pnorms]), "org": np.array([f.origin for f in bfs]),
            "pw": np.array([f.powers for f in bfs], dtype=np.int32)}


def _reference_batches(job):
    """Runs the compiled reference contraction on random unique function
    quartets for `seconds` of compute; returns (quartets done, seconds).
    Module-level so that spawned worker processes can run it."""
    import ctypes

    import numpy as np
    path, method, arrays, seconds, seed = job
    lib = ctypes.CDLL(path)
    PD, PI = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    batch = getattr(lib, "pqref_%s_contr_batch" % method)
    batch.argtypes = [ctypes.c_int, PI, PI, PD, PD, PD, PD, PI, PD]
    rng = np.random.default_rng(seed)
    a_ = arrays
    n = len(a_["nprim"])
    done, t_used, m = 0, 0.0, 2000
    while t_used < seconds:
        q = rng.integers(0, n, size=(m, 4))
        a = np.ascontiguousarray(a_["nprim"][q])
        b = np.ascontiguousarray(a_["poff"][q])
        o = np.ascontiguousarray(a_["org"][q].reshape(m, 12))
        p = np.ascontiguousarray(a_["pw"][q].reshape(m, 12))
        out = np.zeros(m)
        t0 = time.perf_counter()
        batch(m, a.ctypes.data_as(PI), b.ctypes.data_as(PI), a_["e"].ctypes.data_as(PD),
              a_["c"].ctypes.data_as(PD), a_["nr"].ctypes.data_as(PD), o.ctypes.data_as(PD), p.ctypes.data_as(PI),
              out.ctypes.data_as(PD))
        t_used += time.perf_counter() - t0
        done += m
    return done, t_used


def cpu_baseline(bfs, seconds=10.0, seed=0, method="crys"):
    """Time the reference's crys.c contr_coulomb (or chgp.c contr_hrr),
    compiled from the reference sources into oracle/_ref, on random unique
    function quartets, on one thread (the reference is single-threaded)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libref_%s.so" % method)
    if not os.path.exists(path):
        return None
    done, t_used = _reference_batches((path, method, _basis_arrays(bfs), seconds, seed))
    fn = "crys.c contr_coulomb" if method == "crys" else "chgp.c contr_hrr"
    return {"value": done / t_used, "unit": "quartets/s", "cores": 1, "kind": "reference",
            "sample": "%d random unique function quartets of the (H2O)_8 6-31G** basis, reference "
                      "%s compiled -O2, %.1f s, 1 thread" % (done, fn, t_used)}


def cpu_baseline_all_cores(bfs, seconds=5.0, seed=0, method="crys", max_workers=16):
    """The same reference contraction in one process per host core (the
    reference keeps its Rys roots in static arrays, so threads of one process
    would race): the summed rate of up to max_workers spawned workers."""
    import multiprocessing as mp
    path = os.path.join(ROOT, "oracle", "_ref", "libref_%s.so" % method)
    if not os.path.exists(path):
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    workers = max(1, min(max_workers, cores))
    arrays = _basis_arrays(bfs)
    jobs = [(path, method, arrays, seconds, seed + 1 + i) for i in range(workers)]
    with mp.get_context("spawn").Pool(workers) as pool:
        res = pool.map(_reference_batches, jobs)
    rate = sum(d / t for d, t in res)
    fn = "crys.c contr_coulomb" if method == "crys" else "chgp.c contr_hrr"
    return {"value": rate, "unit": "quartets/s", "cores": workers, "kind": "reference",
            "sample": "%d random unique function quartets of the (H2O)_8 6-31G** basis, reference %s compiled "
                      "-O2, %d processes x %.1f s" % (sum(d for d, _ in res), fn, workers, seconds)}


def grad_flops(la, lb, lc, ld):
    """Operation count of the two-electron derivative kernel (FMA = 2 flops),
    the gradient counterpart of the F_prim / F_dig convention: per primitive
    quartet, 30 for the pair setup, and per Rys root (nr = floor((L+1)/2) + 1)
    60 for the root, the 1-D recurrences and transfers of the class one
    higher on A, B and C, (la+1 lb+1 | lc+1 ld), and 51 per Cartesian
    integral (the three yz/xz/xy products, then for each of the centres A, B,
    C and each direction the shifted-pair derivative and its accumulation);
    per contracted quartet 10 per integral for the density factor
    16 D_ab D_cd - 4 (D_ac D_bd + D_ad D_bc).  Returns (per primitive
    quartet, per contracted quartet)."""
    ncart = [1, 3, 6, 10]
    L = la + lb + lc + ld
    nr = (L + 1) // 2 + 1
    a1, b1, c1 = la + 1, lb + 1, lc + 1
    n, m = a1 + b1, c1 + ld
    sb = sum(n - j + 1 for j in range(1, b1 + 1))
    sd = sum(m - l + 1 for l in range(1, ld + 1))
    nc = ncart[la] * ncart[lb] * ncart[lc] * ncart[ld]
    rec = 4 * (n + 1) * (m + 1) + 2 * (m + 1) * sb + 2 * (a1 + 1) * (b1 + 1) * sd
    return 30.0 + nr * (60.0 + 3.0 * rec + 51.0 * nc), 10.0 * nc


def grad_kernel_name(cls):
    return "k_eri_grad<%d, %d, %d, %d>" % tuple(cls)


def recorded(name, key, field):
    """`field` of the entry recorded for `key` (kernel and workload) in
    profiles/<name> from rocprofv3 counter runs, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        entry = json.load(f).get(key)
    return None if entry is None else entry[field]


def recorded_traffic(key):
    """HBM bytes per launch (profiles/traffic.json), or None."""
    return recorded("traffic.json", key, "bytes_per_launch")


def kernel_name(cls, cap, nbf):
    """Kernel template a class's FP64 J/K tiles run in: bra runs (k_eri_group)
    for classes of at most 108 components when the run cap is nonzero."""
    ncart = [1, 3, 6]
    ncomp = ncart[cls[0]] * ncart[cls[1]] * ncart[cls[2]] * ncart[cls[3]]
    if cap is None:
        cap = 16 if nbf >= 300 else 0
    tmpl = "k_eri_group" if ncomp <= 108 and cap > 0 else "k_eri_lane"
    return "%s<%d, %d, %d, %d, double, 0>" % (tmpl, *cls)


def apply_tuning(eng, D, name, args, precision=0, build_tune=True):
    """Sets the engine's bra-run caps for workload `name` (see --tune);
    returns the 21 x 2 cap table in use, or None for the defaults.
    build_tune=False skips the whole-build refinement (per-class autotune
    only), for workloads whose builds are too long to time many times."""
    from pyquante_amd import tuning
    if args.tune == "none":
        return None
    caps = tuning.load(name) if args.tune == "table" else None
    if caps is None:
        caps = eng.autotune(D, reps=10)
        if build_tune and not args.no_build_tune:
            caps, ms = eng.autotune_build(D, caps, precision=precision)
            print("bench: whole-build tuning of %s: %.3f ms/build" % (name, ms), file=sys.stderr, flush=True)
        if args.write_tuning:
            tuning.store(name, caps, nbf=eng.nbf, precision="mixed" if precision else "fp64")
    else:
        eng.set_group_caps(caps)
    return caps


def _timed(fn, steps):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def extra_workloads(args):
    """The workloads reported beside the headline figure (outside its timed
    region): (H2O)_16 FP64 against mixed precision, forces on (H2O)_8, the
    two-density pass, MP2, the batched workload and the HGP path; --only
    selects a subset."""
    import torch

    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.force import scf
    from pyquante_amd.molecule import water_cluster

    out = {}
    only = set(args.only.split(",")) if args.only else None

    def want(name):
        return only is None or name in only

    if want("h2o16"):
        out.update(h2o16_workload(args))
    if not any(want(w) for w in ("forces", "multi", "det", "mp2", "hgp", "batch")):
        return out
    # (H2O)_8: converged RHF for the forces, the two-density pass, MP2 and HGP
    mol = water_cluster((2, 2, 2), seed=args.seed)
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    eng.build(1e-12, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e8, orbe8, C8, D, W = scf(eng, mol, ConvCriteria=1e-9, graph=True)
    torch.cuda.synchronize()
    out["h2o8_rhf_scf_graph_s"] = time.perf_counter() - t0
    if want("forces"):
        out.update(forces_workload(args, eng, D, W))
    if want("multi"):
        # two densities in one pass (the unrestricted J/K) against two builds
        Db_ = (0.9 * D + 0.1 * torch.eye(D.shape[0], dtype=D.dtype, device=D.device)).contiguous()
        t2 = _timed(lambda: eng.fock_jk_multi([D, Db_]), args.steps)
        t1 = _timed(lambda: (eng.fock_jk(D), eng.fock_jk(Db_)), args.steps)
        out.update({"h2o8_two_density_ms": t2 * 1e3, "h2o8_two_single_builds_ms": t1 * 1e3})
        print("bench: (H2O)_8 J/K of two densities %.2f ms in one pass, %.2f ms as two builds" % (t2 * 1e3, t1 * 1e3),
              file=sys.stderr, flush=True)
    if want("det"):
        out.update(deterministic_workload(args, eng, D))
    out.update(extra_h2o8_stored(args, mol, eng, e8, orbe8, C8, want))
    return out


def deterministic_workload(args, eng, D):
    """(H2O)_8: the graph-replayed deterministic build (integer fixed-point
    accumulation) against the FP64-atomic build with the same caps: times,
    the largest J/K difference, and whether repeated replays are
    bit-identical."""
    import torch

    from pyquante_amd import tuning
    caps = tuning.load("h2o8_631ss_fp64")
    if caps is not None:
        eng.set_group_caps(caps)
    g = eng.capture_fock_jk(D)
    t_atomic = _timed(lambda: g(D), args.steps)
    Ja, Ka = [x.clone() for x in g(D)]
    eng.set_deterministic(True)
    g = eng.capture_fock_jk(D)
    t_det = _timed(lambda: g(D), args.steps)
    Jd, Kd = [x.clone() for x in g(D)]
    same = True
    for _ in range(3):
        J, K = g(D)
        same = same and bool(torch.equal(J, Jd)) and bool(torch.equal(K, Kd))
    eng.set_deterministic(False)
    dmax = max(float((Jd - Ja).abs().max()), float((Kd - Ka).abs().max()))
    print("bench: (H2O)_8 deterministic build %.3f ms against %.3f ms with FP64 atomics (x%.2f); max|dJ,dK| %.1e; "
          "replays bit-identical: %s" % (t_det * 1e3, t_atomic * 1e3, t_det / t_atomic, dmax, same),
          file=sys.stderr, flush=True)
    return {"h2o8_deterministic_ms": t_det * 1e3, "h2o8_atomic_ms": t_atomic * 1e3,
            "h2o8_deterministic_max_diff": dmax, "h2o8_deterministic_bit_identical": same}


def h2o16_workload(args):
    """(H2O)_16 6-31G**: graph-replayed Fock builds in FP64 and in mixed
    precision (quartets with Schwarz bound < 1e-4 in FP32), with the RHF
    energies of both."""
    from pyquante_amd import _abi
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.force import scf
    from pyquante_amd.molecule import water_cluster
    out = {}
    mol = water_cluster((2, 2, 4), seed=args.seed)
    eng = Engine(BasisSet(mol, "6-31g**"), mol)
    eng.build(1e-12, 0.0)
    e64, _, _, D64, _ = scf(eng, mol, ConvCriteria=1e-9)
    apply_tuning(eng, D64, "h2o16_631ss_fp64", args)
    t64 = _timed(eng.capture_fock_jk(D64), args.steps)
    st64 = eng.stats()
    eng.build(1e-12, 1e-4)
    st32 = eng.stats()
    emx, _, _, _, _ = scf(eng, mol, ConvCriteria=1e-9, precision=_abi.PQ_PREC_MIXED)
    apply_tuning(eng, D64, "h2o16_631ss_mixed", args, precision=1)
    tmx = _timed(eng.capture_fock_jk(D64, precision=_abi.PQ_PREC_MIXED), args.steps)
    out.update({"h2o16_nbf": eng.nbf, "h2o16_fp64_ms_per_fock_build": t64 * 1e3,
                "h2o16_mixed_ms_per_fock_build": tmx * 1e3, "h2o16_mixed_speedup": t64 / tmx,
                "h2o16_fp64_quartets_per_s": st64["fquartets_fp64"] / t64,
                "h2o16_fp32_quartet_fraction": st32["fquartets_fp32"] / max(1, st32["fquartets_fp32"] +
                                                                               st32["fquartets_fp64"]),
                "h2o16_energy_fp64_eh": e64, "h2o16_energy_mixed_eh": emx,
                "h2o16_mixed_energy_error_eh": abs(emx - e64)})
    print("bench: (H2O)_16 nbf=%d FP64 %.2f ms, mixed %.2f ms (x%.2f), |dE| = %.2e Eh"
          % (eng.nbf, t64 * 1e3, tmx * 1e3, t64 / tmx, abs(emx - e64)), file=sys.stderr, flush=True)
    return out


def forces_workload(args, eng, D, W):
    """Forces per second on (H2O)_8 from the converged density, and the
    roofline of the dominant derivative kernel."""
    out = {}
    tg = _timed(lambda: eng.gradient(D, W), args.steps)
    out.update({"h2o8_ms_per_gradient": tg * 1e3, "h2o8_forces_per_s": 1.0 / tg})
    print("bench: (H2O)_8 gradient %.2f ms (%.1f forces/s)" % (tg * 1e3, 1.0 / tg), file=sys.stderr, flush=True)
    # gradient roofline: the dominant derivative kernel's operation count
    # over its event-timed launch (HIP events on its own stream, the kernels
    # running concurrently as in the timed gradients)
    import numpy as _np
    acc = _np.zeros(49)
    for _ in range(10):
        acc += _np.array(eng.gradient_timed(D, W)[1])
    acc /= 10
    cls = eng.class_stats()
    flops = []
    for c in cls:
        fp, fc = grad_flops(*c["class"])
        flops.append(fp * (c["prim_quartets_fp64"] + c["prim_quartets_fp32"]) +
                     fc * (c["quartets_fp64"] + c["quartets_fp32"]))
    kt = [acc[2 * i] + acc[2 * i + 1] for i in range(21)]
    k = int(_np.argmax(kt))
    ach = flops[k] / (kt[k] * 1e-3) / 1e12
    gname = grad_kernel_name(cls[k]["class"])
    out["roofline_forces"] = {
        "bound": "valu", "kernel": "%s, (%s%s|%s%s) derivative integrals contracted with D" %
        (gname, *("spd"[x] for x in cls[k]["class"])),
        "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
        "launch_ms": kt[k], "flops_per_launch": flops[k],
        "traffic": recorded_traffic("%s h2o8_631ss_grad" % gname),
        "all_kernels_tflops": sum(flops) / (acc[48] * 1e-3) / 1e12,
        "gradient_event_ms": acc[48]}
    print("bench: (H2O)_8 gradient dominant kernel %s %.3f ms, %.2f TFLOP/s (%.3f of peak); all %.2f TFLOP/s"
          % (gname, kt[k], ach, ach / FP64_PEAK_TFLOPS, out["roofline_forces"]["all_kernels_tflops"]),
          file=sys.stderr, flush=True)
    return out


def extra_h2o8_stored(args, mol, eng, e8, orbe8, C8, want):
    """MP2 on (H2O)_8, the batched workload and the HGP path."""
    import numpy as np
    import torch
    out = {}
    # MP2 on (H2O)_8: stored ERIs, row-block unpacking and device GEMMs
    if want("mp2") and not args.no_mp2:
        from pyquante_amd.Ints import Integrals
        from pyquante_amd.MP import MP2
        nocc = mol.get_closedopen()[0]
        nbf = eng.nbf
        orbs = C8.cpu().numpy() if torch.is_tensor(C8) else np.asarray(C8)
        oe = orbe8.cpu().numpy() if torch.is_tensor(orbe8) else np.asarray(orbe8)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ints = Integrals(None, None, engine=eng)
        e2 = MP2(ints, orbs, oe, nocc, nbf - nocc)
        torch.cuda.synchronize()
        tm = time.perf_counter() - t0
        nv = nbf - nocc
        gflop = (2.0 * nocc * nbf ** 4 + 2.0 * nocc ** 2 * nbf ** 3 + 2.0 * nocc ** 2 * nbf ** 2 * nv +
                 2.0 * nocc ** 2 * nv ** 2 * nbf) / 1e9
        out.update({"h2o8_mp2_s": tm, "h2o8_mp2_correlation_eh": e2, "h2o8_rhf_plus_mp2_eh": e8 + e2,
                    "h2o8_mp2_transform_gflop": gflop})
        print("bench: (H2O)_8 MP2 %.3f s (E2 = %.8f Eh, transform %.0f GFLOP)" % (tm, e2, gflop),
              file=sys.stderr, flush=True)
        del ints
    if want("batch") and args.batch_molecules > 0:
        out.update(batch_workload(args))
    # HGP path: every unique function quartet of (H2O)_8 by the GPU HGP
    # kernel, against the Rys kernels' array; RHF through the stored HGP
    # integrals against the direct Rys RHF
    if want("hgp") and not args.no_hgp:
        from pyquante_amd.hartree_fock import rhf
        rys = eng.get2ints("rys")
        hgp = eng.get2ints("hgp", 0)  # first call: kernel loading
        del hgp
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hgp = eng.get2ints("hgp", 0)
        torch.cuda.synchronize()
        th = time.perf_counter() - t0
        dmax = float((hgp - rys).abs().max())
        del hgp
        # the per-function-quartet kernel with chgp's Boys tolerance (the
        # reference's own evaluation order, one function quartet per lane)
        t0 = time.perf_counter()
        hgpf = eng.get2ints("hgp", 1)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t0
        dmax_f = float((hgpf - rys).abs().max())
        del hgpf, rys
        # the same unscreened quartets in the Rys operation count (the
        # HGP path is reported against the Rys-equivalent count)
        eng.build(0.0, 0.0)
        hgp_flops = eng.stats()["flops_fp64"]
        eng.build(1e-12, 0.0)
        from pyquante_amd import Ints
        e_r, _, _ = rhf(mol, basis_data="6-31g**", ConvCriteria=1e-9, MaxIter=60, DoAveraging=True)
        compat = Ints.hgp_compat_boys
        try:
            Ints.hgp_compat_boys = False  # exact Boys values: the two algorithms must agree
            e_h, _, _ = rhf(mol, basis_data="6-31g**", ConvCriteria=1e-9, MaxIter=60, DoAveraging=True,
                            method="hgp")
            Ints.hgp_compat_boys = True   # chgp.c's EPS = 3e-7 Boys function, as the reference computes
            e_hn, _, _ = rhf(mol, basis_data="6-31g**", ConvCriteria=1e-9, MaxIter=60, DoAveraging=True,
                             method="hgp")
        finally:
            Ints.hgp_compat_boys = compat
        out.update({"h2o8_hgp_get2ints_s": th, "h2o8_hgp_quartets_per_s": eng_len(eng) / th,
                    "h2o8_hgp_per_function_kernel_s": tf, "h2o8_hgp_per_function_vs_rys_max_abs": dmax_f,
                    "h2o8_hgp_rys_equivalent_tflops": hgp_flops / th / 1e12,
                    "h2o8_hgp_vs_rys_max_abs": dmax, "h2o8_rhf_energy_rys_eh": e_r, "h2o8_rhf_energy_hgp_eh": e_h,
                    "h2o8_rhf_hgp_minus_rys_eh": e_h - e_r, "h2o8_rhf_hgp_reference_boys_minus_rys_eh": e_hn - e_r})
        if not args.no_cpu:
            from pyquante_amd.basis import BasisSet as _BS
            cpu_h = cpu_baseline(_BS(mol, "6-31g**").bfs, seconds=3.0, method="chgp")
            if cpu_h:
                out["h2o8_hgp_cpu_reference_quartets_per_s"] = cpu_h["value"]
        print("bench: (H2O)_8 HGP get2ints %.3f s by shell quartets (%.2e quartets/s, max |HGP - Rys| %.1e), "
              "%.2f s by function quartets with chgp's Boys (max |d| %.1e); RHF HGP - Rys %.1e Eh"
              % (th, eng_len(eng) / th, dmax, tf, dmax_f, e_h - e_r), file=sys.stderr, flush=True)
    return out


def batch_workload(args):
    """The batched workload: seeded water clusters (H2O)_n, n uniform in 1..12,
    all in one engine (one launch per class over every molecule), RHF for all
    of them at once in FP64 and in mixed precision, forces of all."""
    import torch

    from pyquante_amd import _abi
    from pyquante_amd.batch import BatchEngine, batch_forces, batch_rhf, seeded_water_batch
    mols = seeded_water_batch(nmol=args.batch_molecules, nmax=12, seed=args.seed)
    be = BatchEngine(mols)
    out = {"batch_molecules": be.nmol, "batch_functions": be.nbf}
    for label, fp32, prec in (("fp64", 0.0, _abi.PQ_PREC_FP64), ("mixed", 1e-4, _abi.PQ_PREC_MIXED)):
        be.build(1e-12, fp32)
        # bra-run caps: the kernels' work does not depend on the density's
        # values, so a random packed density serves for live tuning
        Dt = torch.rand(be.mshape, dtype=torch.float64, device="cuda")
        apply_tuning(be, Dt, "batch%d_631ss_%s" % (be.nmol, label), args, precision=prec, build_tune=False)
        del Dt
        st = be.stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E, orbe, C, D, W, it = batch_rhf(be, ConvCriteria=1e-8, precision=prec)
        torch.cuda.synchronize()
        tscf = time.perf_counter() - t0
        tb = _timed(be.capture_fock_jk(D, precision=prec), 3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        F = batch_forces(be, D, W)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t0
        fq = st["fquartets_fp64"] + st["fquartets_fp32"]
        out.update({"batch_%s_fock_build_ms" % label: tb * 1e3, "batch_%s_quartets_per_s" % label: fq / tb,
                    "batch_%s_scf_s" % label: tscf, "batch_%s_scf_iterations" % label: it,
                    "batch_%s_forces_ms" % label: tf * 1e3})
        if label == "fp64":
            E64, F64 = E, F
            out["batch_function_quartets"] = fq
        else:
            out["batch_mixed_max_energy_error_eh"] = float((E - E64).abs().max())
            out["batch_mixed_max_force_error"] = max(float((a - b).abs().max()) for a, b in zip(F, F64))
            out["batch_mixed_fock_speedup"] = out["batch_fp64_fock_build_ms"] / out["batch_mixed_fock_build_ms"]
        print("bench: batch of %d molecules (%d functions) %s: Fock build %.1f ms, SCF %.1f s (%d it), forces %.0f ms"
              % (be.nmol, be.nbf, label, tb * 1e3, tscf, it, tf * 1e3), file=sys.stderr, flush=True)
    print("bench: batch mixed vs FP64: Fock build x%.2f, max |dE| %.1e Eh, max |dF| %.1e"
          % (out["batch_mixed_fock_speedup"], out["batch_mixed_max_energy_error_eh"], out["batch_mixed_max_force_error"]),
          file=sys.stderr, flush=True)
    return out


def eng_len(eng):
    npair = eng.nbf * (eng.nbf + 1) // 2
    return npair * (npair + 1) // 2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the (H2O)_16, forces and HGP workloads")
    ap.add_argument("--no-hgp", action="store_true", help="skip the HGP workload")
    ap.add_argument("--no-mp2", action="store_true", help="skip the MP2 workload")
    ap.add_argument("--only", default="",
                    help="comma-separated subset of the extra workloads to run (h2o16, forces, multi, det, mp2, "
                         "batch, hgp); default all")
    ap.add_argument("--batch-molecules", type=int, default=512,
                    help="molecules of the batched workload (0 skips it)")
    ap.add_argument("--streams", type=int, default=4, help="engine streams for the class kernels (0-4)")
    ap.add_argument("--no-isolated", action="store_true",
                    help="skip timing the class kernels one at a time (keeps a kernel-trace profile's averages to "
                         "the concurrent launches)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device copy (reference for rocprofv3 FETCH/WRITE_SIZE)")
    ap.add_argument("--classes", action="store_true", help="print per-class kernel times")
    ap.add_argument("--tune", choices=["table", "live", "none"], default="table",
                    help="bra-run caps: the stored autotuner table for the workload (falling back to a live "
                         "autotune when none is stored), a live autotune, or the defaults")
    ap.add_argument("--write-tuning", action="store_true", help="store live autotune results in the tuning table")
    ap.add_argument("--no-build-tune", action="store_true",
                    help="with --tune live: keep the per-class autotune result (no whole-build refinement)")
    ap.add_argument("--eager", action="store_true", help="time eager pq_fock_jk calls instead of graph replays")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the Fock build would shard its class work lists per GPU and all-reduce J/K)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from pyquante_amd import _abi
    from pyquante_amd.basis import BasisSet
    from pyquante_amd.engine import Engine
    from pyquante_amd.hartree_fock import chol_orth, geigh, mkdens
    from pyquante_amd.molecule import water_cluster

    torch.cuda.set_device(0)
    mol = water_cluster((2, 2, 2), seed=args.seed)
    bs = BasisSet(mol, "6-31g**")
    eng = Engine(bs, mol)
    eng.set_streams(args.streams)
    eng.build(1e-12, 0.0)
    st = eng.stats()
    npair = eng.nbf * (eng.nbf + 1) // 2
    unscreened = npair * (npair + 1) // 2
    # a representative density: a few SCF iterations from the core guess
    S, T, V = eng.get1ints()
    h = T + V
    X = chol_orth(S)
    _, C = geigh(h, S, X)
    nocc = mol.get_closedopen()[0]
    for _ in range(3):
        D = mkdens(C, 0, nocc)
        J, K = eng.fock_jk(D)
        _, C = geigh(h + 2 * J - K, S, X)
    D = mkdens(C, 0, nocc).contiguous()
    print("bench: (H2O)_8 6-31G** nbf=%d, %d shell quartets, %d function quartets (unscreened %d)"
          % (eng.nbf, st["quartets_fp64"], st["fquartets_fp64"], unscreened), file=sys.stderr, flush=True)
    caps = apply_tuning(eng, D, "h2o8_631ss_fp64", args)
    if caps is not None:
        print("bench: bra-run caps " + " ".join(
            "(%s%s|%s%s)=%d" % (*("spd"[x] for x in cc["class"]), caps[ci][0])
            for ci, cc in enumerate(eng.class_stats()) if cc["quartets_fp64"]), file=sys.stderr, flush=True)
    J0, K0 = eng.fock_jk(D)
    step = (lambda: eng.fock_jk(D)) if args.eager else eng.capture_fock_jk(D)
    for _ in range(max(1, args.warmup)):
        Jg, Kg = step()
    torch.cuda.synchronize()
    graph_err = max(float((Jg - J0).abs().max()), float((Kg - K0).abs().max()))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.fock_jk(D)
    torch.cuda.synchronize()
    dt_eager = (time.perf_counter() - t0) / args.steps
    # per-class kernel times: HIP events on each kernel's launch stream, same
    # concurrent configuration as the timed builds
    cls = eng.class_stats()
    nt = 10
    acc = np.zeros(43)
    for _ in range(nt):
        _, _, ms = eng.fock_jk_timed(D)
        acc += np.array(ms)
    acc /= nt
    if args.classes:
        for ci, cc in enumerate(cls):
            if cc["quartets_fp64"] == 0:
                continue
            t = acc[2 * ci] * 1e-3
            print("  %s quartets %9d prim %11d tiles %7d  %.3f ms  %7.3f TFLOP/s  %6.1f Gquartet/s"
                  % ("(%s%s|%s%s)" % tuple("spd"[x] for x in cc["class"]), cc["quartets_fp64"],
                     cc["prim_quartets_fp64"], cc["tiles_fp64"], t * 1e3,
                     cc["flops_fp64"] / t / 1e12 if t > 0 else 0,
                     cc["fquartets_fp64"] / t / 1e9 if t > 0 else 0), file=sys.stderr)
    # the same kernels one at a time on the caller's stream: each launch with
    # the chip to itself, beside its duration inside the concurrent build
    iso = None
    if not args.no_isolated:
        eng.set_streams(0)
        iso = np.zeros(43)
        for _ in range(5):
            _, _, ms = eng.fock_jk_timed(D)
            iso += np.array(ms)
        iso /= 5
        eng.set_streams(args.streams)
    k = int(np.argmax(acc[:42:2]))
    kname = kernel_name(cls[k]["class"], caps[k][0] if caps else None, eng.nbf)
    traffic = recorded_traffic("%s h2o8_631ss_fp64" % kname)
    # FP64 VALU flops per launch counted by rocprofv3 (64 lanes x (2 FMA +
    # ADD + MUL)), over the same event-timed launch duration
    hw_flops = recorded("valu.json", "%s h2o8_631ss_fp64" % kname, "fp64_flops_per_launch")
    c = cls[k]
    ktime = acc[2 * k] * 1e-3
    achieved = c["flops_fp64"] / ktime / 1e12
    lname = "(%s%s|%s%s)" % tuple("spd"[x] for x in c["class"])
    total_flops = st["flops_fp64"]
    print("bench: %.3f ms/build (%s; eager %.3f ms, replay vs eager max|d| %.1e); dominant kernel %s %.3f ms, "
          "%.2f TFLOP/s; all kernels %.2f TFLOP/s"
          % (dt * 1e3, "eager" if args.eager else "graph replay", dt_eager * 1e3, graph_err, lname, ktime * 1e3,
             achieved, total_flops / (acc[42] * 1e-3) / 1e12), file=sys.stderr, flush=True)
    if iso is not None:
        print("bench: one class at a time: dominant kernel %.3f ms alone (%.3f of FP64 peak), all classes %.3f ms"
              % (iso[2 * k], c["flops_fp64"] / (iso[2 * k] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, iso[42]),
              file=sys.stderr, flush=True)
    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    extra = {}
    if not args.no_extra:
        extra = extra_workloads(args)
    cpu = cpu_all = None
    if not args.no_cpu:
        cpu = cpu_baseline(bs.bfs, seconds=args.cpu_seconds)
        cpu_all = cpu_baseline_all_cores(bs.bfs)
    fq = st["fquartets_fp64"]
    if cpu is not None:
        # integrals only: the screened quartets of one build at the sampled rate
        cpu["fock_build_s_estimate"] = fq / cpu["value"]
    line = {
        "metric": "ERI quartets/s + s per Fock build, mixed vs FP64 at <=1e-6 Eh; forces/s",
        "value": fq / dt,
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded water cluster (numpy default_rng(%d)), published 6-31G** basis" % args.seed,
        "config": {"workload": "(H2O)_8 6-31G** RHF Fock build (J+K), 200 Cartesian functions, "
                               "Schwarz 1e-12, FP64", "nbf": eng.nbf, "shell_quartets": st["quartets_fp64"],
                   "function_quartets": fq, "unscreened_function_quartets": unscreened,
                   "s_per_fock_build": dt, "timing": "eager" if args.eager else "HIP graph replay",
                   "s_per_fock_build_eager": dt_eager, "bra_run_caps": args.tune,
                   "primitive_screening": "on: primitive pairs whose Schwarz factor times the pair's "
                                          "largest is below 1e-4 x the Schwarz cut are skipped",
                   "counting": "value counts the unique function quartets that survive Schwarz screening; "
                               "cpu_baseline counts random unique function quartets, unscreened",
                   **extra},
        "roofline": {"bound": "valu", "compute": "FP64 vector ALU (Rys recurrences are not matrix products, so "
                                                 "no MFMA; 78.6 TFLOP/s is the MI355X FP64 vector peak)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "%s, %s FP64 J/K" % (kname, lname), "launch_ms": ktime * 1e3,
                     "all_kernels_tflops": total_flops / (acc[42] * 1e-3) / 1e12,
                     "counted_fp64_tflops": hw_flops / ktime / 1e12 if hw_flops else None,
                     "isolated": None if iso is None else {
                         "launch_ms": iso[2 * k],
                         "frac": c["flops_fp64"] / (iso[2 * k] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                         "serialized_build_ms": iso[42],
                         "note": "the same kernel launched alone (class kernels one at a time on one stream); "
                                 "frac above is its duration inside the concurrent build"}},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    if "roofline_forces" in line["config"]:
        line["roofline_forces"] = line["config"].pop("roofline_forces")
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
