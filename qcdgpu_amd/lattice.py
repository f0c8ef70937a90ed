"""Host-side mirror of QCDGPU's model driver for the update/measurement path.

`Lattice` plays the role of model_CL::model (SU(N)/suncl/suncl.h:61-467) for
this path: it holds the parameters the reference reads from init.dat and its
hardcoded defaults (SU(N)/QCDGPU.cpp:98-235), owns one device lattice in
libqcdhip.so, and runs the reference's loop (suncl.cpp:5212-5505): NAV
thermalisation sweeps, then ITER measurements each after NITER sweeps, the
first measurement on the start configuration.  Normalisations follow
model::lattice_analysis (suncl.cpp:1328-1612).
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import check, lib

_START = {"hot": _abi.START_HOT, "cold": _abi.START_COLD, "gid": _abi.START_GID, "random": _abi.START_RANDOM}


class Lattice:
    def __init__(
        self,
        sun=3,
        dims=(32, 32, 32, 8),
        beta=5.9,
        precision="fp32",
        storage_reals=None,
        nhit=10,
        prng="XOR128",
        ranlux_luxury=3,
        stream="reference",
        randseries=54,
        max_measurements=1024,
        prng_instances=0,
    ):
        """prng_instances: the reference's PRNG_instances (0: one instance
        per link of a half-step, the fused per-site streams; otherwise a
        multiple of 64 and the reference's PRN buffer is produced before
        every heatbath half-step)."""
        p = _abi.Params()
        p.sun = sun
        for i in range(4):
            p.n[i] = int(dims[i])
        p.precision = _abi.FP32 if precision == "fp32" else _abi.FP64
        p.storage_reals = storage_reals or (18 if sun == 3 else 4)
        p.beta = float(beta)
        p.nhit = int(nhit)
        p.nhitpar = 1
        p.prng = {"XOR128": _abi.PRNG_XOR128, "RANLUX": _abi.PRNG_RANLUX, "PM": _abi.PRNG_PM}[prng]
        p.ranlux_luxury = int(ranlux_luxury)
        p.stream_mode = _abi.STREAM_REFERENCE if stream == "reference" else _abi.STREAM_FAST
        p.randseries = int(randseries)
        p.max_measurements = int(max_measurements)
        p.prng_instances = int(prng_instances)
        self.params = p
        self.sun = sun
        self.dims = tuple(int(d) for d in dims)
        self.volume = int(np.prod(self.dims))
        self._h = ctypes.c_void_p()
        self._owned = True
        check(lib().qcd_lattice_create(ctypes.byref(p), ctypes.byref(self._h)))

    @classmethod
    def from_params(cls, params):
        """A lattice created from a filled _abi.Params (e.g. a .qcg header)."""
        self = cls.__new__(cls)
        p = _abi.Params()
        ctypes.pointer(p)[0] = params
        self.params = p
        self.sun = p.sun
        self.dims = tuple(int(p.n[k]) for k in range(4))
        self.volume = int(np.prod(self.dims))
        self._h = ctypes.c_void_p()
        self._owned = True
        check(lib().qcd_lattice_create(ctypes.byref(p), ctypes.byref(self._h)))
        return self

    @classmethod
    def from_qcg(cls, path):
        """model::lattice_init + lattice_load_state for INIT = 0: a lattice
        with the parameters of the state file's header, loaded from it.
        Returns (lattice, run counters)."""
        params, _, _ = read_qcg_header(path)
        lat = cls.from_params(params)
        return lat, lat.load_qcg(path)

    @classmethod
    def _view(cls, handle, params):
        """A non-owning view of a lattice owned elsewhere (a batch chain)."""
        self = cls.__new__(cls)
        self.params = params
        self.sun = params.sun
        self.dims = tuple(int(params.n[k]) for k in range(4))
        self.volume = int(np.prod(self.dims))
        self._h = handle
        self._owned = False
        return self

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if self._h and self._owned:
            check(lib().qcd_lattice_destroy(self._h))
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # -- starts, PRNG, host copies -----------------------------------------
    def init(self, start="cold"):
        """start: "cold", "gid", "hot" (Haar links from the streams; call
        prng_init first) or "random" (the reference's hot start, which seeds
        the streams itself, suncl.cpp:5234-5248)."""
        check(lib().qcd_lattice_init(self._h, _START[start]))

    def prng_init(self):
        check(lib().qcd_prng_init(self._h))

    def link_shape(self):
        return (self.volume, 4, self.sun, self.sun, 2)

    def export_links(self):
        a = np.zeros(self.link_shape())
        check(lib().qcd_lattice_export(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return a

    def import_links(self, links):
        a = np.ascontiguousarray(links, dtype=np.float64).reshape(self.link_shape())
        check(lib().qcd_lattice_import(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    # -- updates --------------------------------------------------------------
    def halfstep(self, parity, mu, kind="hb"):
        check(lib().qcd_halfstep(self._h, parity, mu, _abi.HEATBATH if kind == "hb" else _abi.OVERRELAX))

    def sweep(self, kind="hb"):
        check(lib().qcd_sweep(self._h, _abi.HEATBATH if kind == "hb" else _abi.OVERRELAX))

    def reunitarize(self):
        check(lib().qcd_reunitarize(self._h))

    def sync(self):
        check(lib().qcd_lattice_sync(self._h))

    # -- measurements ---------------------------------------------------------
    def measure(self, plaquette=True, polyakov=True, wilson=False, polyakov_slices=False, fmunu=None,
                action_slices=False):
        """fmunu: None, "H" (planes xy, xz, yz) or "E" (xt, yt, zt); it
        and action_slices imply the plaquette."""
        if fmunu not in (None, "H", "E"):
            raise ValueError("fmunu must be None, 'H' or 'E'")
        mask = (
            (_abi.MEAS_PLAQUETTE if plaquette else 0)
            | (_abi.MEAS_POLYAKOV if polyakov else 0)
            | (_abi.MEAS_WILSON if wilson else 0)
            | (_abi.MEAS_POLYAKOV_SLICES if polyakov_slices else 0)
            | {None: 0, "H": _abi.MEAS_FMUNU, "E": _abi.MEAS_F0MU}[fmunu]
            | (_abi.MEAS_ACTION_SLICES if action_slices else 0)
        )
        check(lib().qcd_measure(self._h, mask))

    def action_slices(self, raw=False):
        """Action per x, y and z slice for every plaquette record: dict
        x/y/z -> array (count, N_d, 2) of (spatial, temporal).  Raw: sums of
        ReTr U_p over the slice's three spatial / three temporal plaquettes
        per site; otherwise the reference's S_diff_s / S_diff_t,
        beta/N (N - sum / (3 * sites in the slice)) (suncl.cpp:1376-1480)."""
        n1, n2, n3, n4 = self.dims
        cap = self.params.max_measurements
        a = np.zeros((cap, n1 + n2 + n3, 2))
        c = ctypes.c_int()
        check(lib().qcd_read_action_slices(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           ctypes.byref(c)))
        a = a[: min(c.value, cap)]
        out = {"x": a[:, :n1], "y": a[:, n1:n1 + n2], "z": a[:, n1 + n2:]}
        if not raw:
            V, N = float(self.volume), float(self.sun)
            bet = self.params.beta / N
            for k, n in zip("xyz", (n1, n2, n3)):
                out[k] = bet * (N - out[k] / (3.0 * V / n))
        return out

    def set_fmunu(self, gen1, gen2):
        """Generators of the F_munu measurement: SU(3) lambda_gen1 (1..4) and
        lambda_gen2 (5..8), default 3 and 8; SU(2) sigma_gen1 (1 or 3) and
        sigma_2 (suncl.cpp:2093-2113)."""
        check(lib().qcd_lattice_set_fmunu(self._h, int(gen1), int(gen2)))

    def fmunu(self, raw=False):
        """F_munu records, one per plaquette measurement: array (count, 2, 3)
        complex, [generator][plane], planes (x,y), (x,z), (y,z) for the H
        field or (x,t), (y,t), (z,t) for the E field.  Raw: sums over sites
        of -i Tr(P G) / 2; otherwise divided by V, as the reference's
        analysis (suncl.cpp:1614-1623)."""
        cap = self.params.max_measurements
        a = np.zeros((cap, 2, 3, 2))
        c = ctypes.c_int()
        check(lib().qcd_read_fmunu(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(c)))
        a = a[: min(c.value, cap)]
        z = a[..., 0] + 1j * a[..., 1]
        return z if raw else z / float(self.volume)

    def polyakov_slices(self, raw=False):
        """Polyakov loop per x, y and z slice for every polyakov_slices
        measurement: dict x/y/z -> complex array (count, N_d).  Raw: sums of
        Tr P over the slice; otherwise divided by N2 N3 N, the denominator
        the reference's analysis uses for all three directions
        (suncl.cpp:1516, 1541, 1566)."""
        n1, n2, n3 = self.dims[:3]
        cap = self.params.max_measurements
        a = np.zeros((cap, n1 + n2 + n3, 2))
        c = ctypes.c_int()
        check(lib().qcd_read_polyakov_slices(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                             ctypes.byref(c)))
        a = a[: min(c.value, cap)]
        z = a[..., 0] + 1j * a[..., 1]
        if not raw:
            z = z / (n2 * n3 * self.sun)
        return {"x": z[:, :n1], "y": z[:, n1:n1 + n2], "z": z[:, n1 + n2:]}

    def set_twist(self, phi, omega=0.0):
        """Twisted boundary (PHI, OMEGA): Y links at x = N1 - 1 are read as
        T U, T = diag(e^{i(phi+omega)/2}, e^{i(omega-phi)/2}, e^{-i omega})
        (SU(2): diag(e^{i phi/2}, e^{-i phi/2})), su3_matrix_memory.cl:57-98."""
        check(lib().qcd_lattice_set_twist(self._h, float(phi), float(omega)))

    def twist(self):
        phi, omega = ctypes.c_double(), ctypes.c_double()
        check(lib().qcd_lattice_get_twist(self._h, ctypes.byref(phi), ctypes.byref(omega)))
        return phi.value, omega.value

    def set_wilson(self, R, T):
        """Extents of the Wilson loop measured with measure(wilson=True):
        R along x, y, z and T along t (reference defaults 7 and 2)."""
        check(lib().qcd_lattice_set_wilson(self._h, int(R), int(T)))

    def wilson_measurements(self, raw=False):
        """Wilson loop records: sums of ReTr W over sites and spatial
        directions, or (default) divided by 3V as the reference's analysis
        (suncl.cpp:1606-1611)."""
        cap = self.params.max_measurements
        w = np.zeros(cap)
        n = ctypes.c_int()
        check(lib().qcd_read_wilson(self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n)))
        w = w[: min(n.value, cap)]
        return w if raw else w / (3.0 * self.volume)

    def reset_measurements(self):
        check(lib().qcd_reset_measurements(self._h))

    def raw_measurements(self):
        cap = self.params.max_measurements
        plq = np.zeros((cap, 2))
        poly = np.zeros((cap, 4))
        n1, n2 = ctypes.c_int(), ctypes.c_int()
        D = ctypes.POINTER(ctypes.c_double)
        check(
            lib().qcd_read_measurements(
                self._h, plq.ctypes.data_as(D), poly.ctypes.data_as(D), ctypes.byref(n1), ctypes.byref(n2)
            )
        )
        return plq[: min(n1.value, cap)], poly[: min(n2.value, cap)]

    def measurements(self):
        """Normalised like model::lattice_analysis (suncl.cpp:1356-1373, 1483-1494):
        Plq_spat/temp = sum / (3 V) (mean ReTr U_p, not divided by N),
        Plq_total = (spat + temp) / 2, Polyakov = sum / (V_s N),
        |P|^2 site average / N^2, |P|^4 site average / N^4."""
        plq, poly = self.raw_measurements()
        V = float(self.volume)
        Vs = V / self.dims[3]
        N = float(self.sun)
        out = {}
        if len(plq):
            out["plq_spat"] = plq[:, 0] / (3 * V)
            out["plq_temp"] = plq[:, 1] / (3 * V)
            out["plq_total"] = 0.5 * (out["plq_spat"] + out["plq_temp"])
        if len(poly):
            out["polyakov_re"] = poly[:, 0] / (Vs * N)
            out["polyakov_im"] = poly[:, 1] / (Vs * N)
            out["polyakov_p2"] = poly[:, 2] / (Vs * N * N)
            out["polyakov_p4"] = poly[:, 3] / (Vs * N ** 4)
        return out

    # -- graphs ---------------------------------------------------------------
    def capture_cycle(self, n_hb=1, n_or=0, plaquette=True, polyakov=True, wilson=False):
        mask = ((_abi.MEAS_PLAQUETTE if plaquette else 0) | (_abi.MEAS_POLYAKOV if polyakov else 0)
                | (_abi.MEAS_WILSON if wilson else 0))
        check(lib().qcd_cycle_capture(self._h, n_hb, n_or, mask))

    def run_cycle(self, n_hb=1, n_or=0, plaquette=True, polyakov=True):
        """The measurement cycle enqueued eagerly (same kernels as the graph)."""
        mask = (_abi.MEAS_PLAQUETTE if plaquette else 0) | (_abi.MEAS_POLYAKOV if polyakov else 0)
        check(lib().qcd_cycle_run(self._h, n_hb, n_or, mask))

    def launch_cycle(self, repeats=1):
        check(lib().qcd_cycle_launch(self._h, repeats))

    def time_cycle(self, repeats):
        ms = ctypes.c_double()
        check(lib().qcd_cycle_time(self._h, repeats, ctypes.byref(ms)))
        return ms.value

    def profile_cycle(self, repeats):
        hb, orr, meas, tot = (ctypes.c_double() for _ in range(4))
        check(
            lib().qcd_cycle_profile(
                self._h, repeats, ctypes.byref(hb), ctypes.byref(orr), ctypes.byref(meas), ctypes.byref(tot)
            )
        )
        return {"hb_us": hb.value, "or_us": orr.value, "meas_us": meas.value, "total_ms": tot.value}

    def tune(self, xcd_map=None, blocks_per_cu=None, update_shape=None, shadow=None, cycles_per_graph=None):
        """Launch-shape knobs of the update kernels (scheduling only; results
        are identical): xcd_map 0/1/2/3, blocks_per_cu cap (0 = none),
        update_shape 3 (three wavefronts per 64 links) or 1; shadow = True
        makes each half-step gather from a copy of the table taken before
        it (a race check: the links must come out the same);
        cycles_per_graph K: capture_cycle also records K cycles as one graph
        with each measurement beside the next cycle's first half-steps."""
        if xcd_map is not None:
            check(lib().qcd_lattice_tune(self._h, _abi.TUNE_XCD_MAP, int(xcd_map)))
        if blocks_per_cu is not None:
            check(lib().qcd_lattice_tune(self._h, _abi.TUNE_BLOCKS_PER_CU, int(blocks_per_cu)))
        if update_shape is not None:
            check(lib().qcd_lattice_tune(self._h, _abi.TUNE_UPDATE_SHAPE, int(update_shape)))
        if shadow is not None:
            check(lib().qcd_lattice_tune(self._h, _abi.TUNE_SHADOW, int(bool(shadow))))
        if cycles_per_graph is not None:
            check(lib().qcd_lattice_tune(self._h, _abi.TUNE_CYCLES_PER_GRAPH, int(cycles_per_graph)))

    def kernel(self, name):
        """Id of the reference kernel `name` bound to this lattice
        (GPU::kernel_init's registry)."""
        i = ctypes.c_int()
        check(lib().qcd_kernel_create(self._h, name.encode(), ctypes.byref(i)))
        return i.value

    def run_kernel(self, kid, wait=True):
        check((lib().qcd_kernel_run if wait else lib().qcd_kernel_enqueue)(self._h, int(kid)))

    def kernel_info(self, kid):
        name = ctypes.create_string_buffer(64)
        n = ctypes.c_uint()
        check(lib().qcd_kernel_info(self._h, int(kid), name, 64, ctypes.byref(n)))
        return name.value.decode(), n.value

    def graph_begin(self):
        check(lib().qcd_graph_begin(self._h))

    def graph_end(self):
        check(lib().qcd_graph_end(self._h))

    def set_slab(self, on=True):
        """x-slab mode: x = 0 and x = N1 - 1 are halo slices that half-steps
        read and never update (one GPU's share of a lattice cut along x)."""
        check(lib().qcd_lattice_set_slab(self._h, 1 if on else 0))

    def slab_face_bytes(self):
        b = ctypes.c_size_t()
        check(lib().qcd_slab_face_bytes(self._h, ctypes.byref(b)))
        return b.value

    def slab_face(self, side, parity, mu, dev_ptr, unpack):
        """Copy the (parity, mu) links of the first (side 0) or last (side 1)
        interior slice into device memory at dev_ptr (unpack False), or
        device memory into the halo on that side (unpack True)."""
        check(lib().qcd_slab_face(self._h, int(side), int(parity), int(mu), ctypes.c_void_p(dev_ptr),
                                  1 if unpack else 0))

    def tuning(self):
        """Knob values the update launches use: xcd_map after its fallbacks
        (with the XCD block's x-values and z-rows), blocks_per_cu,
        update_shape."""
        v = ctypes.c_int()
        det = (ctypes.c_int * 2)()
        check(lib().qcd_lattice_get_tune(self._h, _abi.TUNE_XCD_MAP, ctypes.byref(v), det))
        out = {"xcd_map": v.value, "xcd_block": (det[0], det[1])}
        for name, knob in (("blocks_per_cu", _abi.TUNE_BLOCKS_PER_CU), ("update_shape", _abi.TUNE_UPDATE_SHAPE),
                           ("shadow", _abi.TUNE_SHADOW)):
            check(lib().qcd_lattice_get_tune(self._h, knob, ctypes.byref(v), None))
            out[name] = v.value
        check(lib().qcd_lattice_get_tune(self._h, _abi.TUNE_CYCLES_PER_GRAPH, ctypes.byref(v), det))
        out["cycles_per_graph"] = v.value
        out["captured_cycles_per_graph"] = det[0]
        return out

    def table_bytes(self):
        b = ctypes.c_size_t()
        check(lib().qcd_lattice_table_bytes(self._h, ctypes.byref(b)))
        return b.value

    # -- checkpoints (model::lattice_save_state / lattice_load_state) ----------
    def prng_counter(self):
        """PRNG::produce calls the reference would have made so far."""
        c = ctypes.c_uint()
        check(lib().qcd_prng_counter(self._h, ctypes.byref(c)))
        return c.value

    def prng_replay(self, counter, start="cold"):
        """Re-seed and advance the streams to production `counter` of a run
        that began with `start` (suncl.cpp:5349-5351)."""
        check(lib().qcd_prng_replay(self._h, int(counter), _START[start] if isinstance(start, str) else int(start)))

    def save_state(self):
        """Exact state image (uint8 array): table, streams, records, counters."""
        n = ctypes.c_size_t()
        check(lib().qcd_lattice_state_bytes(self._h, ctypes.byref(n)))
        a = np.empty(n.value, dtype=np.uint8)
        check(lib().qcd_lattice_state_save(self._h, a.ctypes.data_as(ctypes.c_void_p), a.size))
        return a

    def load_state(self, image):
        a = np.ascontiguousarray(np.frombuffer(image, dtype=np.uint8))
        check(lib().qcd_lattice_state_load(self._h, a.ctypes.data_as(ctypes.c_void_p), a.size))

    def write_measurements(self, plq=None, poly=None, wilson=None):
        """Replace the raw records (plq: (n, 2), poly: (n, 4), wilson: (n,))."""
        D = ctypes.POINTER(ctypes.c_double)
        arrs, counts = [], []
        for x, w in ((plq, 2), (poly, 4), (wilson, 1)):
            if x is None:
                arrs.append(None)
                counts.append(0)
            else:
                a = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, w)
                arrs.append(a)
                counts.append(a.shape[0])
        ptr = [a.ctypes.data_as(D) if a is not None and a.size else None for a in arrs]
        check(lib().qcd_write_measurements(self._h, ptr[0], ptr[1], ptr[2], *counts))

    def save_qcg(self, path, **run):
        """Write the reference's .qcg state file (suncl.cpp:1857-1993) plus
        the exact state image.  `run` sets qcd_run_info fields (init, start,
        nav, nav_counter, niter, iter, iter_counter, wilson_r, wilson_t,
        pl_level, get_actions, get_plaquettes, get_wilson, fmunu_index1,
        fmunu_index2, get_fmunu, get_f0mu); by default ITER is the record
        capacity and ITER_counter the number of plaquette records.  With
        get_fmunu or get_f0mu the file carries the F_munu records."""
        plq, _ = self.raw_measurements()
        info = _abi.RunInfo()
        defaults = dict(init=1, start="cold", nav=0, nav_counter=0, niter=1, iter=self.params.max_measurements,
                        iter_counter=len(plq), wilson_r=7, wilson_t=2, pl_level=3, get_actions=1,
                        get_plaquettes=1, get_wilson=0, fmunu_index1=3, fmunu_index2=8 if self.sun == 3 else 2,
                        get_fmunu=0, get_f0mu=0)
        unknown = set(run) - set(defaults) - {"prng_counter"}
        if unknown:
            raise ValueError("unknown run fields %s" % sorted(unknown))
        defaults.update(run)
        if isinstance(defaults["start"], str):
            defaults["start"] = _START[defaults["start"]]
        for k, v in defaults.items():
            setattr(info, k, int(v))
        check(lib().qcd_qcg_write(self._h, str(path).encode(), ctypes.byref(info)))

    def load_qcg(self, path):
        """Load a .qcg file into this lattice; returns the header's run counters."""
        info = _abi.RunInfo()
        check(lib().qcd_qcg_read(self._h, str(path).encode(), ctypes.byref(info)))
        return info.as_dict()

    # -- the reference run loop -------------------------------------------------
    def simulate(self, nav=0, iters=2, niter=1, n_or=0, start="cold", wilson=False, slices=False, fmunu=None,
                 action_slices=False):
        """model::lattice_simulate: start, measure, NAV sweeps, then ITER-1
        measurements each after NITER sweeps (suncl.cpp:5236-5483).  Each sweep
        is one heatbath sweep plus `n_or` overrelaxation sweeps; wilson=True
        adds the Wilson loop (set_wilson) to every measurement, slices=True the
        Polyakov loop per slice (PL_level > 2), fmunu="H"/"E" the field
        strength (get_Fmunu / get_F0mu), action_slices=True the action per
        slice (get_actions_diff)."""
        if start == "hot":
            self.prng_init()   # the Haar start draws from the streams
            self.init(start)
        else:
            self.init(start)
            if start not in ("gid", "random"):   # the random start seeds the streams itself
                self.prng_init()
        self.reset_measurements()
        self.measure(wilson=wilson, polyakov_slices=slices, fmunu=fmunu, action_slices=action_slices)
        for _ in range(nav):
            self.sweep("hb")
            for _ in range(n_or):
                self.sweep("or")
        for _ in range(iters - 1):
            for _ in range(niter):
                self.sweep("hb")
                for _ in range(n_or):
                    self.sweep("or")
            self.measure(wilson=wilson, polyakov_slices=slices, fmunu=fmunu, action_slices=action_slices)
        self.sync()
        return self.measurements()


def read_qcg_header(path):
    """model::lattice_load_bin_header (suncl.cpp:1995-2041): the lattice
    parameters and run counters of a .qcg file, and whether it carries the
    exact state image.  Reads the file only (no device)."""
    p = _abi.Params()
    info = _abi.RunInfo()
    ex = ctypes.c_int()
    check(lib().qcd_qcg_read_header(str(path).encode(), ctypes.byref(p), ctypes.byref(info), ctypes.byref(ex)))
    return p, info.as_dict(), bool(ex.value)


def device_info():
    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    mem = ctypes.c_size_t()
    check(lib().qcd_device_info(name, 256, ctypes.byref(cus), ctypes.byref(mem)))
    return {"name": name.value.decode(), "compute_units": cus.value, "global_mem": mem.value}


def prng_produce(prng, table, samples, luxury=3):
    """Reference PRNG::produce buffer semantics on the device: returns
    quads[samples, n, 4]; `table` is advanced in place.  XOR128 / PM: uint32
    [n, 4]; RANLUX: float32 [7, n, 4] (six float4 of lagged values and
    (23, 9, block counter, carry) per instance)."""
    code = {"XOR128": _abi.PRNG_XOR128, "PM": _abi.PRNG_PM, "RANLUX": _abi.PRNG_RANLUX}[prng]
    if prng == "RANLUX":
        t = np.ascontiguousarray(table, dtype=np.float32)
        n = t.shape[1]
    else:
        t = np.ascontiguousarray(table, dtype=np.uint32)
        n = t.shape[0]
    q = np.zeros((samples, n, 4), dtype=np.float32)
    fn = lib().qcd_prng_produce_host
    check(fn(code, int(luxury), n, samples, t.ctypes.data_as(ctypes.c_void_p),
             q.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    table[...] = t
    return q
