"""The one-GPU share of the x-slab decomposition (the reference's BIGLAT
path: every device holds Nx + 2 x-slices, suncl.cpp:2422, and UpdateEdges,
suncl.cpp:4005-4062, refreshes the edges): a periodic lattice cut into two
x-slabs with one halo slice on each side.  After every half-step each slab
packs the updated (parity, direction) links of its two boundary slices into
a device buffer and the other slab unpacks them into its halos -- what RCCL
send / recv would move between x-neighbours.  The slabs' interiors must then
equal the periodic lattice updated whole, bit for bit, through heatbath
(XOR128 reference streams, the slab's streams taken from the same sites) and
overrelaxation sweeps."""
import ctypes

import numpy as np
import pytest

import qcdgpu_amd as Q
from qcdgpu_amd import _abi

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case]", msg, flush=True)


def _hip():
    """hipMalloc / hipFree of the HIP runtime libqcdhip.so itself uses
    (symbols looked up through the library's handle search its
    dependencies; another copy of the runtime may be loaded in the process,
    e.g. PyTorch's)."""
    h = _abi.lib()
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMalloc.restype = ctypes.c_int
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipFree.restype = ctypes.c_int
    return h


def _rng_slice(img, lat):
    """The generator states in a state image: (streams, 4) uint32."""
    tb = lat.table_bytes()
    rb = lat.params.n[1] * lat.params.n[2] * lat.params.n[3] * lat.params.n[0] // 2 * 16
    return 136 + tb, 136 + tb + rb


@pytest.mark.parametrize("prec,storage,sun", [("fp64", 18, 3), ("fp32", 18, 3), ("fp32", 12, 3), ("fp64", 4, 2)])
def test_two_slabs_equal_periodic_lattice(prec, storage, sun):
    _case("slab %s-%d su%d" % (prec, storage, sun))
    N1, rest = 8, (8, 4, 4)
    ns = N1 // 2                                   # interior x-slices per slab
    dims = (N1,) + rest
    sdims = (ns + 2,) + rest
    beta = 5.9 if sun == 3 else 2.3
    kw = dict(sun=sun, precision=prec, storage_reals=storage if sun == 3 else None, beta=beta, randseries=17,
              max_measurements=2)
    hip = _hip()
    with Q.Lattice(dims=dims, **kw) as P, Q.Lattice(dims=sdims, **kw) as A, Q.Lattice(dims=sdims, **kw) as B:
        P.prng_init()
        P.init("hot")
        P.sweep("hb")
        full = P.export_links().reshape((N1, -1) + P.link_shape()[1:])
        Vx = full.shape[1]
        img = P.save_state()
        a0, a1 = _rng_slice(img, P)
        rng = img[a0:a1].view(np.uint32).reshape(-1, 4)
        hn = Vx // 2
        slabs = []
        for lat, x0 in ((A, 0), (B, ns)):          # interior global x0 .. x0 + ns - 1
            xs = [(x0 - 1 + k) % N1 for k in range(ns + 2)]
            lat.prng_init()
            lat.init("cold")
            lat.import_links(full[xs].reshape(lat.link_shape()))
            simg = lat.save_state()
            b0, b1 = _rng_slice(simg, lat)
            srng = simg[b0:b1].view(np.uint32).reshape(-1, 4)
            for k, x in enumerate(xs):
                srng[k * hn:(k + 1) * hn] = rng[x * hn:(x + 1) * hn]
            lat.load_state(simg)
            lat.set_slab(True)
            flip = (x0 - 1) % 2                   # slab parity = global parity ^ flip
            slabs.append((lat, x0, flip))
        nbytes = A.slab_face_bytes()
        bufs = [ctypes.c_void_p() for _ in range(4)]
        for b in bufs:
            assert hip.hipMalloc(ctypes.byref(b), nbytes) == 0
        try:
            for kind in ("hb", "or", "hb"):
                for p in (1, 0):
                    for mu in range(4):
                        P.halfstep(p, mu, kind)
                        for lat, _, flip in slabs:
                            lat.halfstep(p ^ flip, mu, kind)
                        # A's high interior -> B's low halo, B's low interior -> A's high halo, and the wrap
                        (lA, _, fA), (lB, _, fB) = slabs
                        lA.slab_face(1, p ^ fA, mu, bufs[0].value, False)
                        lB.slab_face(0, p ^ fB, mu, bufs[1].value, False)
                        lA.slab_face(0, p ^ fA, mu, bufs[2].value, False)
                        lB.slab_face(1, p ^ fB, mu, bufs[3].value, False)
                        lA.sync()
                        lB.sync()
                        lB.slab_face(0, p ^ fB, mu, bufs[0].value, True)
                        lA.slab_face(1, p ^ fA, mu, bufs[1].value, True)
                        lB.slab_face(1, p ^ fB, mu, bufs[2].value, True)
                        lA.slab_face(0, p ^ fA, mu, bufs[3].value, True)
                full = P.export_links().reshape((N1, -1) + P.link_shape()[1:])
                for lat, x0, _ in slabs:
                    got = lat.export_links().reshape((ns + 2, -1) + P.link_shape()[1:])
                    xs = [(x0 - 1 + k) % N1 for k in range(ns + 2)]
                    # interior and (freshly exchanged) halos alike
                    assert np.array_equal(got, full[xs]), (kind, x0, float(np.abs(got - full[xs]).max()))
        finally:
            for b in bufs:
                hip.hipFree(b)


def test_slab_refusals():
    _case("slab refusals")
    with Q.Lattice(sun=3, dims=(8, 4, 4, 2), beta=5.9, max_measurements=2) as lat:   # N2 N3 N4 / 2 = 16
        with pytest.raises(Q.QCDError, match="divisible by 64"):
            lat.set_slab(True)
    with Q.Lattice(sun=3, dims=(6, 8, 4, 4), beta=5.9, max_measurements=2) as lat:
        lat.set_slab(True)
        with pytest.raises(Q.QCDError, match="twisted"):
            lat.set_twist(1.0)
