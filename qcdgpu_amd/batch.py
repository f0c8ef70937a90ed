"""Batched ensembles: many independent chains (own extents, beta, random
series) updated together by libqcdhip.so, one launch per half-step and
bucket of equal spatial extent (qcd_batch_* in include/qcd_hip.h)."""
import ctypes

import numpy as np

from . import _abi
from ._abi import check, lib
from .lattice import Lattice


class Batch:
    def __init__(
        self,
        chains,
        sun=3,
        precision="fp32",
        storage_reals=None,
        nhit=10,
        prng="RANLUX",
        ranlux_luxury=3,
        stream="reference",
        max_measurements=1024,
    ):
        """chains: sequence of (dims, beta, randseries)."""
        p = _abi.Params()
        p.sun = sun
        p.precision = _abi.FP32 if precision == "fp32" else _abi.FP64
        p.storage_reals = storage_reals or (18 if sun == 3 else 4)
        p.nhit = int(nhit)
        p.nhitpar = 1
        p.prng = {"XOR128": _abi.PRNG_XOR128, "RANLUX": _abi.PRNG_RANLUX, "PM": _abi.PRNG_PM}[prng]
        p.ranlux_luxury = int(ranlux_luxury)
        p.stream_mode = _abi.STREAM_REFERENCE if stream == "reference" else _abi.STREAM_FAST
        p.max_measurements = int(max_measurements)
        self.params = p
        arr = (_abi.ChainDesc * len(chains))()
        self.chain_params = []
        for k, (dims, beta, seed) in enumerate(chains):
            for j in range(4):
                arr[k].n[j] = int(dims[j])
            arr[k].beta = float(beta)
            arr[k].randseries = int(seed)
            q = _abi.Params.from_buffer_copy(p)
            for j in range(4):
                q.n[j] = int(dims[j])
            q.beta = float(beta)
            q.randseries = int(seed)
            self.chain_params.append(q)
        self._h = ctypes.c_void_p()
        check(lib().qcd_batch_create(ctypes.byref(p), len(chains), arr, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            check(lib().qcd_batch_destroy(self._h))
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

    def info(self):
        nc, nb = ctypes.c_int(), ctypes.c_int()
        check(lib().qcd_batch_info(self._h, ctypes.byref(nc), ctypes.byref(nb)))
        return {"chains": nc.value, "buckets": nb.value}

    def init(self, start="hot"):
        check(lib().qcd_batch_init(self._h, {"hot": _abi.START_HOT, "cold": _abi.START_COLD}[start]))

    def sweep(self, kind="hb"):
        check(lib().qcd_batch_sweep(self._h, _abi.HEATBATH if kind == "hb" else _abi.OVERRELAX))

    @staticmethod
    def _mask(plaquette, polyakov):
        return (_abi.MEAS_PLAQUETTE if plaquette else 0) | (_abi.MEAS_POLYAKOV if polyakov else 0)

    def run_cycle(self, n_hb=1, n_or=0, plaquette=True, polyakov=True):
        check(lib().qcd_batch_cycle_run(self._h, n_hb, n_or, self._mask(plaquette, polyakov)))

    def capture_cycle(self, n_hb=1, n_or=0, plaquette=True, polyakov=True):
        check(lib().qcd_batch_cycle_capture(self._h, n_hb, n_or, self._mask(plaquette, polyakov)))

    def launch_cycle(self, repeats=1):
        check(lib().qcd_batch_cycle_launch(self._h, repeats))

    def time_cycle(self, repeats):
        ms = ctypes.c_double()
        check(lib().qcd_batch_cycle_time(self._h, repeats, ctypes.byref(ms)))
        return ms.value

    def sync(self):
        check(lib().qcd_batch_sync(self._h))

    def autotune(self, repeats=4):
        """Picks the update-kernel shape and the unit placement over the
        XCDs per bucket and update kind; returns tuning()."""
        nb = self.info()["buckets"]
        out = (ctypes.c_int * (2 * nb))()
        check(lib().qcd_batch_autotune(self._h, int(repeats), out))
        return self.tuning()

    def set_tuning(self, bucket, shape=(3, 3), placement=(0, 0)):
        check(lib().qcd_batch_set_tuning(self._h, int(bucket), int(shape[0]), int(shape[1]), int(placement[0]),
                                         int(placement[1])))

    def set_merge(self, on=True):
        """One launch per (parity, mu) and kernel shape for all buckets (on,
        the default) or one per bucket; the chains' results are the same."""
        check(lib().qcd_batch_set_merge(self._h, 1 if on else 0))

    def tuning(self):
        """Per bucket: {"n_s", "chains", "shape": (heatbath, overrelaxation),
        "placement": (heatbath, overrelaxation)}."""
        res = []
        for k in range(self.info()["buckets"]):
            o = (ctypes.c_int * 6)()
            check(lib().qcd_batch_get_tuning(self._h, k, o))
            res.append({"n_s": o[4], "chains": o[5], "shape": (o[0], o[1]), "placement": (o[2], o[3])})
        return res

    def chain(self, index):
        """Chain `index` as a non-owning Lattice view (export, records)."""
        h = ctypes.c_void_p()
        check(lib().qcd_batch_chain(self._h, int(index), ctypes.byref(h)))
        return Lattice._view(h, self.chain_params[index])


def polyakov_susceptibility(measurements, dims, sun):
    """chi_P = V_s (<|P|^2> - <|P|>^2) of the volume-averaged Polyakov loop
    P = sum_x Tr P(x) / (V_s N) of each configuration (not the site average
    of |Tr P(x)|^2 that the reference's P2 record holds)."""
    P = measurements["polyakov_re"] + 1j * measurements["polyakov_im"]
    a = np.abs(P)
    vs = dims[0] * dims[1] * dims[2]
    return vs * (np.mean(a ** 2) - np.mean(a) ** 2)
