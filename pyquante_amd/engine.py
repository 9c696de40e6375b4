"""Engine: one basis on one molecule, resident on the GPU.

Holds the libpqhip engine handle (shell data, shell-pair SoA in HBM and the
Schwarz-sorted class work lists).  Density, J, K, S, T, V and forces are
torch tensors on the current HIP device; the engine reads and writes them
through their device pointers.
"""
import ctypes

import numpy as np
import torch

from . import _abi
from ._abi import check, lib, pd, pi
from .basis import sym2l


def shells_of(bfs):
    """Shell list from a BasisSet, or rebuilt from a flat list of CGBFs by
    grouping runs of functions with the same centre and primitives that form
    a complete Cartesian set in PyQuante order."""
    if hasattr(bfs, "shells"):
        return [(sh.L, sh.exps, sh.coefs, sh.atid, sh.origin) for sh in bfs.shells]
    from .basis import sym2powerlist
    order = {L: sym2powerlist[s] for s, L in sym2l.items()}
    out = []
    i = 0
    bl = list(bfs)
    while i < len(bl):
        f = bl[i]
        L = sum(f.powers)
        comps = order.get(L)
        if comps is None:
            raise ValueError("only s, p, d functions are supported")
        group = bl[i:i + len(comps)]
        if (len(group) != len(comps) or [tuple(g.powers) for g in group] != comps or
                any(g.origin != f.origin or g.pexps != f.pexps or g.pcoefs != f.pcoefs for g in group)):
            raise ValueError("basis functions %d.. do not form a complete shell in PyQuante order" % i)
        _check_norms(group, i)
        out.append((L, list(f.pexps), list(f.pcoefs), getattr(f, "atid", 0), f.origin))
        i += len(comps)
    return out


def _check_norms(group, first):
    """The engine derives every normalisation itself: the primitive norm of
    Src/lib/primitive-gto.c:76-83 per Cartesian component, and no contracted
    renormalisation (norm = 1, CGBF.py:34).  A caller whose CGBFs carry other
    values would silently get other integrals, so refuse them."""
    from .basis import prim_norm
    for k, g in enumerate(group):
        norm = getattr(g, "norm", 1.0)
        if abs(norm - 1.0) > 1e-12:
            raise ValueError("basis function %d has norm %r: the engine uses norm = 1 for every "
                             "contracted function (no renormalisation)" % (first + k, norm))
        pnorms = getattr(g, "pnorms", None)
        if pnorms is None:
            continue
        for a, n in zip(g.pexps, pnorms):
            ref = prim_norm(float(a), tuple(g.powers))
            if abs(n - ref) > 1e-10 * abs(ref):
                raise ValueError("basis function %d has primitive norm %r for exponent %r, the engine "
                                 "computes %r (per-component Cartesian norm)" % (first + k, n, a, ref))


def _device():
    _abi.require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Engine(object):
    """Engine(bfs, atoms): bfs is a BasisSet (or CGBF list), atoms a Molecule
    or a list of Atom."""

    def __init__(self, bfs, atoms, atom_mol=None, shells=None, shell_atoms=None):
        self.device = _device()
        atoms = list(atoms)
        self.natom = len(atoms)
        self.atom_z = np.array([a.atno for a in atoms], dtype=np.float64)
        xyz = np.array([a.pos() for a in atoms], dtype=np.float64)
        shells = shells_of(bfs) if shells is None else shells
        # map each shell to an atom by its centre (batches give the map: their
        # molecules may overlap in space)
        self._shell_atom = []
        if shell_atoms is not None:
            self._shell_atom = [int(k) for k in shell_atoms]
        else:
            for (L, e, c, atid, origin) in shells:
                d = np.sum((xyz - np.array(origin)) ** 2, axis=1)
                k = int(np.argmin(d))
                if d[k] > 1e-20:
                    raise ValueError("shell centre does not coincide with an atom")
                self._shell_atom.append(k)
        self.shell_l = np.array([s[0] for s in shells], dtype=np.int32)
        nprim = [len(s[1]) for s in shells]
        self.shell_nprim = np.array(nprim, dtype=np.int32)
        self.shell_poff = np.concatenate([[0], np.cumsum(nprim)[:-1]]).astype(np.int32)
        self.shell_atom = np.array(self._shell_atom, dtype=np.int32)
        self.exps = np.array([e for s in shells for e in s[1]], dtype=np.float64)
        self.coefs = np.array([c for s in shells for c in s[2]], dtype=np.float64)
        h = ctypes.c_void_p()
        am = None if atom_mol is None else np.ascontiguousarray(np.asarray(atom_mol, dtype=np.int32))
        check(lib.pq_engine_create_batch(len(shells), pi(self.shell_l), pi(self.shell_nprim), pi(self.shell_poff),
                                         pi(self.shell_atom), pd(self.exps), pd(self.coefs), self.natom,
                                         pd(np.ascontiguousarray(xyz.reshape(-1))), pd(self.atom_z),
                                         None if am is None else pi(am), ctypes.byref(h)))
        self.handle = h
        n = ctypes.c_int()
        check(lib.pq_engine_nbf(h, ctypes.byref(n)))
        self.nbf = n.value
        nm, ms = ctypes.c_int(), ctypes.c_int64()
        check(lib.pq_engine_layout(h, ctypes.byref(nm), ctypes.byref(ms), None, None, None, None))
        self.nmol, self.msize = nm.value, ms.value
        # matrices: n x n for one molecule, packed per-molecule blocks otherwise
        self.mshape = (self.nbf, self.nbf) if self.nmol == 1 else (self.msize,)
        self.xyz = xyz
        self.cut = None
        self.fp32cut = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            lib.pq_engine_destroy(h)
            self.handle = None

    # ------------------------------------------------------------------
    def set_geometry(self, xyz):
        xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(self.natom, 3))
        check(lib.pq_engine_set_geometry(self.handle, pd(xyz.reshape(-1))))
        self.xyz = xyz
        self.cut = self.fp32cut = None

    def build(self, schwarz_cut=1e-12, fp32_cut=0.0):
        """Shell pairs, Schwarz bounds and work lists (pq_build_pairs)."""
        check(lib.pq_build_pairs(self.handle, float(schwarz_cut), float(fp32_cut)))
        self.cut = schwarz_cut
        self.fp32cut = fp32_cut
        return self

    def generation(self):
        """Work-list generation (pq_engine_generation): changes whenever the
        device arrays a captured Fock build reads are replaced."""
        g = ctypes.c_uint64()
        check(lib.pq_engine_generation(self.handle, ctypes.byref(g)))
        return g.value

    def set_streams(self, n):
        """Spread class kernels over n engine streams (0 = caller's stream only)."""
        check(lib.pq_engine_set_streams(self.handle, int(n)))

    def set_fused(self, on):
        """One launch for the light classes' FP64 bra runs instead of one per
        class (pq_engine_set_fused)."""
        check(lib.pq_engine_set_fused(self.handle, 1 if on else 0))

    def set_deterministic(self, on):
        """Bit-reproducible J/K (pq_engine_set_deterministic): exact integer
        fixed-point accumulation instead of FP64 atomics, so every build of
        the same density with the same caps gives the same bits, whatever the
        execution order (streams, graph replay); one or two densities, FP64
        or mixed precision."""
        check(lib.pq_engine_set_deterministic(self.handle, 1 if on else 0))
        self.deterministic = bool(on)

    def set_lds_limit(self, nbytes=0):
        """Dynamic LDS limit of the J/K launches (pq_engine_set_lds_limit;
        0 = the device's); returns the device's per-workgroup LDS bytes."""
        dev = ctypes.c_int()
        check(lib.pq_engine_set_lds_limit(self.handle, int(nbytes), ctypes.byref(dev)))
        return dev.value

    def set_prim_screen(self, rel):
        """Primitive screening of the J/K and gradient work relative to the
        Schwarz cut (pq_engine_set_prim_screen; 0 = every primitive); applies
        from the next build()."""
        check(lib.pq_engine_set_prim_screen(self.handle, float(rel)))

    def set_group_cap(self, cap):
        """Tiles per bra run for every class (0 = none, -1 = default); a run
        length may be or-ed with _abi.PQ_CAP_STAGE_D (bra rows of D in LDS)."""
        check(lib.pq_engine_set_group_cap(self.handle, int(cap)))

    def set_group_caps(self, caps):
        """Per-class bra-run caps, a 21 x 2 list (class, FP64/FP32 tiles) as
        returned by autotune; -1 keeps the default."""
        flat = np.ascontiguousarray(np.asarray(caps, dtype=np.int32).reshape(42))
        check(lib.pq_engine_set_group_caps(self.handle, pi(flat)))

    def autotune(self, D, reps=3):
        """Times each class's J/K kernel on density D for the candidate bra
        run lengths, with and without staged D rows, and keeps the fastest
        (pq_engine_autotune); returns the
        chosen cap per (class, precision) as a 21 x 2 list."""
        if self.cut is None:
            self.build()
        D = self._dtensor(D)
        caps = (ctypes.c_int * 42)()
        check(lib.pq_engine_autotune(self.handle, ctypes.c_void_p(D.data_ptr()), int(reps), _stream(), caps))
        torch.cuda.synchronize(self.device)
        return [[caps[2 * c], caps[2 * c + 1]] for c in range(21)]

    def autotune_build(self, D, caps=None, precision=_abi.PQ_PREC_FP64, reps=5, steps=10, min_gain=0.003):
        """Whole-build refinement of a cap table (per-class autotune result if
        caps is None): for each class without a component split, in order of
        primitive work, tries every candidate cap while the other classes keep
        theirs, timing graph-replayed builds (median of `reps` runs of `steps`
        builds), and keeps a change when the build gets at least `min_gain`
        faster.  The class kernels of a build run concurrently, so the best cap
        of a class alone is not always the best inside the build.  Returns
        (caps, ms per build)."""
        import statistics
        import time
        if self.cut is None:
            self.build()
        prec = 1 if precision == _abi.PQ_PREC_MIXED else 0
        caps = [list(c) for c in (caps if caps is not None else self.autotune(D, reps=10))]
        S = _abi.PQ_CAP_STAGE_D
        cands = [0, 4, 8, 16, 32, 4 | S, 8 | S, 16 | S, 32 | S]
        ncart = [1, 3, 6]

        def timed(table):
            self.set_group_caps(table)
            g = self.capture_fock_jk(D, precision)
            for _ in range(3):
                g(D)
            torch.cuda.synchronize(self.device)
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                for _ in range(steps):
                    g(D)
                torch.cuda.synchronize(self.device)
                ts.append((time.perf_counter() - t0) / steps)
            return statistics.median(ts)

        best = timed(caps)
        cls = self.class_stats()
        key = "prim_quartets_fp32" if prec else "prim_quartets_fp64"
        order = sorted(range(len(cls)), key=lambda c: -cls[c][key])
        for c in order:
            la, lb, lc, ld = cls[c]["class"]
            if cls[c][key] == 0 or ncart[la] * ncart[lb] * ncart[lc] * ncart[ld] > 81:
                continue
            for cap in cands:
                if cap == caps[c][prec]:
                    continue
                trial = [list(x) for x in caps]
                trial[c][prec] = cap
                t = timed(trial)
                if t < best * (1.0 - min_gain):
                    best, caps = t, trial
        self.set_group_caps(caps)
        return caps, best * 1e3

    def stats(self):
        counts = (ctypes.c_uint64 * 6)()
        flops = (ctypes.c_double * 2)()
        check(lib.pq_engine_stats(self.handle, counts, flops))
        return {"quartets_fp64": counts[0], "quartets_fp32": counts[1], "prim_quartets_fp64": counts[2],
                "prim_quartets_fp32": counts[3], "fquartets_fp64": counts[4], "fquartets_fp32": counts[5],
                "flops_fp64": flops[0], "flops_fp32": flops[1]}

    def class_stats(self):
        """Per-class counts and algorithmic flops (pq_engine_class_stats)."""
        out = []
        for c in range(21):
            l = (ctypes.c_int * 4)()
            counts = (ctypes.c_uint64 * 8)()
            flops = (ctypes.c_double * 2)()
            check(lib.pq_engine_class_stats(self.handle, c, l, counts, flops))
            out.append({"class": tuple(l), "quartets_fp64": counts[0], "quartets_fp32": counts[1],
                        "prim_quartets_fp64": counts[2], "prim_quartets_fp32": counts[3],
                        "fquartets_fp64": counts[4], "fquartets_fp32": counts[5],
                        "tiles_fp64": counts[6], "tiles_fp32": counts[7],
                        "flops_fp64": flops[0], "flops_fp32": flops[1]})
        return out

    def fock_jk_timed(self, D, precision=_abi.PQ_PREC_FP64):
        """fock_jk with per-class kernel times in ms (HIP events on the
        current stream); returns (J, K, ms[43])."""
        D = self._dtensor(D)
        J = torch.empty_like(D)
        K = torch.empty_like(D)
        ms = (ctypes.c_float * 43)()
        check(lib.pq_fock_jk_timed(self.handle, ctypes.c_void_p(D.data_ptr()), ctypes.c_void_p(J.data_ptr()),
                                   ctypes.c_void_p(K.data_ptr()), int(precision), _stream(), ms))
        return J, K, list(ms)

    def _dtensor(self, D):
        if not torch.is_tensor(D):
            D = torch.as_tensor(np.asarray(D, dtype=np.float64))
        D = D.to(device=self.device, dtype=torch.float64).contiguous()
        if tuple(D.shape) != self.mshape:
            raise ValueError("density must have shape %s" % (self.mshape,))
        return D

    def fock_jk(self, D, want_j=True, want_k=True, precision=_abi.PQ_PREC_FP64):
        """(J, K) for a symmetric density D in direct mode (pq_fock_jk)."""
        if self.cut is None:
            self.build()
        D = self._dtensor(D)
        J = torch.empty_like(D) if want_j else None
        K = torch.empty_like(D) if want_k else None
        check(lib.pq_fock_jk(self.handle, ctypes.c_void_p(D.data_ptr()),
                             ctypes.c_void_p(J.data_ptr() if J is not None else 0),
                             ctypes.c_void_p(K.data_ptr() if K is not None else 0), int(precision),
                             _stream()))
        return J, K

    def fock_jk_multi(self, Ds):
        """[(J, K) for D in Ds] for one or two symmetric densities from one
        pass over the integrals (pq_fock_jk_multi, FP64)."""
        if self.cut is None:
            self.build()
        if self.fp32cut:
            self.build(self.cut, 0.0)
        Ds = [self._dtensor(D) for D in Ds]
        if not 1 <= len(Ds) <= 2:
            raise ValueError("one or two densities")
        Js = [torch.empty_like(D) for D in Ds]
        Ks = [torch.empty_like(D) for D in Ds]
        arr = ctypes.c_void_p * len(Ds)
        check(lib.pq_fock_jk_multi(self.handle, len(Ds), arr(*[D.data_ptr() for D in Ds]),
                                   arr(*[J.data_ptr() for J in Js]), arr(*[K.data_ptr() for K in Ks]),
                                   _abi.PQ_PREC_FP64, _stream()))
        return list(zip(Js, Ks))

    def shell_quartets(self, quartets):
        """Blocks of the shell quartets [(a, b, c, d), ...] (engine shell
        indices) in the libint layout (pq_shell_quartet_blocks): a list of
        numpy arrays of shape (na, nb, nc, nd)."""
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.int32).reshape(-1, 4))
        if len(q) == 0:
            return []
        nf = np.array([(L + 1) * (L + 2) // 2 for L in self.shell_l], dtype=np.int64)
        if q.min() < 0 or q.max() >= len(nf):
            raise ValueError("shell index out of range")
        sizes = np.prod(nf[q], axis=1)
        out = torch.empty(int(sizes.sum()), dtype=torch.float64, device=self.device)
        check(lib.pq_shell_quartet_blocks(self.handle, len(q), pi(q.reshape(-1)), ctypes.c_void_p(out.data_ptr()),
                                          ctypes.c_uint64(out.numel()), _stream()))
        host = out.cpu().numpy()
        blocks, o = [], 0
        for t, sz in enumerate(sizes):
            blocks.append(host[o:o + sz].reshape(tuple(nf[q[t]])))
            o += sz
        return blocks

    def capture_fock_jk(self, D, precision=_abi.PQ_PREC_FP64):
        """Captures one direct Fock build into a HIP graph (FockGraph)."""
        return FockGraph(self, D, precision)

    def gradient(self, D, W):
        """dE/dR (natom x 3, hartree/bohr) of the RHF energy at density D and
        energy-weighted density W (pq_forces)."""
        if self.cut is None:
            self.build()
        D = self._dtensor(D)
        W = self._dtensor(W)
        g = torch.empty((self.natom, 3), dtype=torch.float64, device=self.device)
        check(lib.pq_forces(self.handle, ctypes.c_void_p(D.data_ptr()), ctypes.c_void_p(W.data_ptr()),
                            ctypes.c_void_p(g.data_ptr()), _stream()))
        return g

    def gradient_uhf(self, Da, Db, Qa, Qb):
        """dE/dR (natom x 3, hartree/bohr) of the UHF energy at spin densities
        Da, Db and energy-weighted densities Qa, Qb (pq_forces_uhf)."""
        if self.cut is None:
            self.build()
        Da, Db = self._dtensor(Da), self._dtensor(Db)
        Qa, Qb = self._dtensor(Qa), self._dtensor(Qb)
        D = (0.5 * (Da + Db)).contiguous()
        W = (0.5 * (Qa + Qb)).contiguous()
        g = torch.empty((self.natom, 3), dtype=torch.float64, device=self.device)
        check(lib.pq_forces_uhf(self.handle, ctypes.c_void_p(Da.data_ptr()), ctypes.c_void_p(Db.data_ptr()),
                                ctypes.c_void_p(D.data_ptr()), ctypes.c_void_p(W.data_ptr()),
                                ctypes.c_void_p(g.data_ptr()), _stream()))
        return g

    def gradient_timed(self, D, W):
        """gradient with per-kernel times in ms (pq_forces_timed): returns
        (dE/dR, ms[49]): ms[2c + prec] the derivative kernel of class c over
        its FP64 / FP32-bucket tiles, ms[42 + p] the one-electron derivative
        kernel of pair class p, ms[48] the whole gradient."""
        if self.cut is None:
            self.build()
        D = self._dtensor(D)
        W = self._dtensor(W)
        g = torch.empty((self.natom, 3), dtype=torch.float64, device=self.device)
        ms = (ctypes.c_float * 49)()
        check(lib.pq_forces_timed(self.handle, ctypes.c_void_p(D.data_ptr()), ctypes.c_void_p(W.data_ptr()),
                                  ctypes.c_void_p(g.data_ptr()), _stream(), ms))
        return g, list(ms)

    def get2ints(self, method="rys", compat_nr=0):
        """The reference's flat unique-ERI array as a device tensor:
        method "rys" (pq_get2ints, shell-quartet Rys kernels) or "hgp"
        (pq_get2ints_hgp; compat_nr = 1 keeps chgp's Boys tolerance)."""
        n = self.nbf
        npair = n * (n + 1) // 2
        length = npair * (npair + 1) // 2
        out = torch.empty(length, dtype=torch.float64, device=self.device)
        if method == "rys":
            check(lib.pq_get2ints(self.handle, ctypes.c_void_p(out.data_ptr()), ctypes.c_uint64(length), _stream()))
        elif method == "hgp":
            check(lib.pq_get2ints_hgp(self.handle, ctypes.c_void_p(out.data_ptr()), ctypes.c_uint64(length),
                                      int(compat_nr), _stream()))
        else:
            raise ValueError("method must be 'rys' or 'hgp'")
        return out

    def get1ints(self, compat_nr=0):
        """(S, T, V) device tensors (pq_get1ints); h = T + V.  compat_nr = 1
        evaluates V with the reference's EPS = 3e-7 Boys function (its
        cints values); 0 with full-precision Boys values, the V whose
        derivative the gradient kernels contract."""
        S = torch.empty(self.mshape, dtype=torch.float64, device=self.device)
        T = torch.empty_like(S)
        V = torch.empty_like(S)
        check(lib.pq_get1ints(self.handle, ctypes.c_void_p(S.data_ptr()), ctypes.c_void_p(T.data_ptr()),
                              ctypes.c_void_p(V.data_ptr()), int(compat_nr), _stream()))
        return S, T, V


def digest_stored(ints, D, want_j=True, want_k=True):
    """J, K from a flat unique-ERI array (device tensor) and density D."""
    n = D.shape[0]
    J = torch.empty_like(D) if want_j else None
    K = torch.empty_like(D) if want_k else None
    check(lib.pq_digest_stored(n, ctypes.c_void_p(ints.data_ptr()), ctypes.c_uint64(ints.numel()),
                               ctypes.c_void_p(D.data_ptr()),
                               ctypes.c_void_p(J.data_ptr() if J is not None else 0),
                               ctypes.c_void_p(K.data_ptr() if K is not None else 0), _stream()))
    return J, K


class StaleGraphError(RuntimeError):
    """A captured Fock build whose engine has replaced its work lists."""


class FockGraph:
    """One pq_fock_jk call captured into a HIP graph on fixed D/J/K buffers.

    Calling it copies a new density into the captured D buffer (when given)
    and replays the graph: the fill, class and symmetrisation kernels of the
    build, with the engine's stream fork/join, are launched as one graph.
    The returned J and K are the graph's own buffers, overwritten by the
    next replay.  The graph holds the device addresses of the engine's work
    lists: build(), set_geometry(), set_group_cap(s)() and autotune() replace
    them, and a call after any of them raises StaleGraphError (checked on
    the host against the engine's work-list generation) instead of replaying.
    """

    def __init__(self, engine, D, precision=_abi.PQ_PREC_FP64):
        self.engine = engine
        self.precision = precision
        self.D = engine._dtensor(D).clone()
        side = torch.cuda.Stream(device=engine.device)
        side.wait_stream(torch.cuda.current_stream(engine.device))
        with torch.cuda.stream(side):
            # first call outside capture: creates the engine's streams/events
            engine.fock_jk(self.D, precision=precision)
        torch.cuda.current_stream(engine.device).wait_stream(side)
        torch.cuda.synchronize(engine.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.J, self.K = engine.fock_jk(self.D, precision=precision)
        torch.cuda.synchronize(engine.device)
        self.generation = engine.generation()

    def __call__(self, D=None):
        if self.engine.generation() != self.generation:
            raise StaleGraphError("the engine's work lists changed since this Fock build was captured; "
                                  "capture it again")
        if D is not None:
            self.D.copy_(self.engine._dtensor(D))
        self.graph.replay()
        return self.J, self.K
