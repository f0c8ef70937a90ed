"""Restatement of the reference's .qcg state-file layout in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the
library's writer and reader, never on the product path.

Sources (SU(N)/...):
  header words          suncl/suncl.cpp:1939-1993 (lattice_make_bin_header),
                        read back at :1995-2041 (lattice_load_bin_header)
  string words          suncl/suncl.cpp:709-716 (convert_str_uint)
  codes                 suncl/suncl.cpp:718-744 (start, precision),
                        random/random.cpp:110-122 (generator), RL_nskip
                        random/random.cpp:66,192-196
  double split          clinterface/clinterface.cpp:325-336 (LOW / HIGH words)
  section order         suncl/suncl.cpp:1889-1901 (lattice_save_state)
  F_munu sections       suncl/suncl.cpp:2266-2298 (x MODEL_energies_size = 7,
                        suncl/suncl.h:221), block order of
                        reduce_measurement_plq_double2
                        (suncl/sun_measurements_cl.cl:1312-1354)
  sizes                 suncl/suncl.cpp:2266-2298 with buffer_size_align
                        (clinterface/clinterface.cpp:1587-1590) at the
                        64-item work-group cap of QCDGPU.cpp:74
  table planes          suncl/su3_matrix_memory.cl:150-176 (lattice_store_3),
                        suncl/su2_matrix_memory.cl:128-145 (lattice_store_2);
                        decode suncl/su3cl.cl:137-163, suncl/su2cl.cl:63-79
"""
import numpy as np

HEAD_WORDS = 64
EXT_MAGIC = b"QCDGPUX1"
GENERATOR = {"XOR128": 1, "RANLUX0": 2, "RANLUX1": 3, "RANLUX2": 4, "RANLUX3": 5, "RANLUX4": 6, "PM": 9}
NSKIP = {"RANLUX0": 24, "RANLUX1": 48, "RANLUX2": 97, "RANLUX3": 223, "RANLUX4": 389}

# header word index of each field, in lattice_make_bin_header's order
FIELDS = [
    "init", "start", "randseries", "generator", "rl_nskip", "prng_counter", "nav", "nav_counter", "niter",
    "iter", "iter_counter", "nhit", "wilson_r", "wilson_t", "precision", "beta_lo", "beta_hi", "phi_lo",
    "phi_hi", "omega_lo", "omega_hi", "pl_level", "fmunu_index1", "fmunu_index2", "get_fmunu", "get_actions",
    "get_plaquettes", "get_wilson", "measurement_size_F", "energies_size", "energies_size_F",
    "polyakov_loop_size", "group", "nd",
]
WORD = {name: 4 + k for k, name in enumerate(FIELDS)}   # sizes follow at 38..45


def string_word(s, offset):
    """Bytes offset..offset+3 of s, the first in the lowest byte."""
    b = s.encode()[offset:offset + 4]
    return sum(c << (8 * i) for i, c in enumerate(b))


def align(x, step=64):
    return ((x - 1) // step + 1) * step


def sizes(group, dims, iters):
    V = int(np.prod(dims))
    row = align(V)
    elements = 12 if group == 3 else 4
    return {
        "V": V,
        "rowsize": row,
        "table_size": align(row * 4 * elements // 4),
        "measurement_size_F": align(-(-align(V) // 64)),
        "energies_size": align(iters),
        "polyakov_loop_size": align(iters),
    }


def make_header(group, dims, beta, precision, generator, f_sections=False, **fields):
    """The 64 header words; `fields` are FIELDS entries (counters, flags).
    f_sections: the partials and plaquette sections carry F_munu (7 x)."""
    w = np.zeros(HEAD_WORDS, dtype=np.uint32)
    w[0], w[1] = string_word("QCDGPU", 0), string_word("QCDGPU", 4)
    w[2] = string_word("1.5", 0)
    s = sizes(group, dims, fields.get("iter", 1))
    vals = dict(
        init=1, start=1, randseries=54, generator=GENERATOR[generator], rl_nskip=NSKIP.get(generator, 24),
        prng_counter=0, nav=0, nav_counter=0, niter=1, iter=1, iter_counter=0, nhit=10, wilson_r=7, wilson_t=2,
        precision={"single": 1, "double": 2}[precision], pl_level=3, get_actions=1, get_plaquettes=1,
        get_wilson=0, measurement_size_F=s["measurement_size_F"] * (7 if f_sections else 1),
        energies_size=s["energies_size"], energies_size_F=s["energies_size"] * (7 if f_sections else 1),
        polyakov_loop_size=s["polyakov_loop_size"], group=group, nd=4, fmunu_index1=3,
        fmunu_index2=8 if group == 3 else 2,
    )
    bits = np.array([beta], dtype=np.float64).view(np.uint32)
    vals["beta_lo"], vals["beta_hi"] = int(bits[0]), int(bits[1])
    vals.update(fields)
    for name, v in vals.items():
        w[WORD[name]] = np.uint32(v)
    for k in range(4):
        w[38 + k] = w[42 + k] = dims[k]
    return w


def parse_header(w):
    w = np.asarray(w, dtype=np.uint32)
    out = {name: int(w[i]) for name, i in WORD.items()}
    out["prefix"] = (int(w[0]), int(w[1]))
    out["version"] = (int(w[2]), int(w[3]))
    out["beta"] = float(np.array([w[WORD["beta_lo"]], w[WORD["beta_hi"]]], dtype=np.uint32).view(np.float64)[0])
    out["full"] = tuple(int(x) for x in w[38:42])
    out["domain"] = tuple(int(x) for x in w[42:46])
    return out


def planes_from_links(links, group, rowsize, table_size):
    """Reference table (table_size, 4) from canonical links (V, 4, N, N, 2):
    entry gindex + (4 * plane + dir) * ROWSIZE."""
    V = links.shape[0]
    t = np.zeros((table_size, 4))
    re, im = links[..., 0], links[..., 1]
    if group == 3:
        planes = [
            np.stack([re[:, :, 0, 0], re[:, :, 0, 1], re[:, :, 0, 2], re[:, :, 1, 2]], axis=-1),
            np.stack([im[:, :, 0, 0], im[:, :, 0, 1], im[:, :, 0, 2], im[:, :, 1, 2]], axis=-1),
            np.stack([re[:, :, 1, 0], re[:, :, 1, 1], im[:, :, 1, 0], im[:, :, 1, 1]], axis=-1),
        ]
    else:
        planes = [np.stack([re[:, :, 0, 0], re[:, :, 0, 1], im[:, :, 0, 0], im[:, :, 0, 1]], axis=-1)]
    for k, pl in enumerate(planes):
        for d in range(4):
            base = (4 * k + d) * rowsize
            t[base:base + V] = pl[:, d]
    return t


def links_from_planes(table, group, V, rowsize):
    """Canonical links from the reference table, rebuilding the third SU(3)
    row as conj(row1 x row2) and the SU(2) second row as (-conj u12, conj u11)."""
    N = group
    L = np.zeros((V, 4, N, N), dtype=np.complex128)
    for d in range(4):
        p0 = table[d * rowsize:d * rowsize + V]
        if N == 3:
            p1 = table[(4 + d) * rowsize:(4 + d) * rowsize + V]
            p2 = table[(8 + d) * rowsize:(8 + d) * rowsize + V]
            u = np.stack([p0[:, 0] + 1j * p1[:, 0], p0[:, 1] + 1j * p1[:, 1], p0[:, 2] + 1j * p1[:, 2]], axis=-1)
            v = np.stack([p2[:, 0] + 1j * p2[:, 2], p2[:, 1] + 1j * p2[:, 3], p0[:, 3] + 1j * p1[:, 3]], axis=-1)
            L[:, d, 0], L[:, d, 1] = u, v
            L[:, d, 2] = np.conj(np.cross(u, v))
        else:
            u11, u12 = p0[:, 0] + 1j * p0[:, 2], p0[:, 1] + 1j * p0[:, 3]
            L[:, d, 0, 0], L[:, d, 0, 1] = u11, u12
            L[:, d, 1, 0], L[:, d, 1, 1] = -np.conj(u12), np.conj(u11)
    return np.stack([L.real, L.imag], axis=-1)


def _has_f(h):
    return h["energies_size_F"] != h["energies_size"]


def encode(header, links, plq=(), poly=(), wilson=(), fmunu=()):
    """A reference-layout file: header, partials (zeros), action sums from
    the plaquette sums, plaquette sums (then, with F sections, the F_munu
    sums: fmunu[rec, generator, plane] complex), Wilson and Polyakov (Re, Im)
    sums, table (float4 for precision word 1, double4 otherwise).  Section
    sizes are the header's words."""
    h = parse_header(header)
    single = h["precision"] == 1
    group, dims = h["group"], h["full"]
    s = sizes(group, dims, h["iter"])
    out = [np.asarray(header, dtype=np.uint32).tobytes(), np.zeros(2 * h["measurement_size_F"]).tobytes()]
    en = h["energies_size"]
    plq = np.asarray(plq, dtype=np.float64).reshape(-1, 2)
    bet, V = h["beta"] / group, float(s["V"])
    act = np.zeros((en, 2))
    act[: len(plq)] = bet * (3.0 * group * V - plq)
    out.append(act.tobytes())
    if h["get_plaquettes"] or h["get_fmunu"] or _has_f(h):
        e = np.zeros((h["energies_size_F"], 2))
        e[: len(plq)] = plq
        f = np.asarray(fmunu, dtype=np.complex128).reshape(-1, 6)
        for b in range(6 if _has_f(h) else 0):   # block 1 + 3 * generator + plane
            e[(1 + b) * en:(1 + b) * en + len(f), 0] = f[:, b].real
            e[(1 + b) * en:(1 + b) * en + len(f), 1] = f[:, b].imag
        out.append(e.tobytes())
    if h["get_wilson"]:
        e = np.zeros(en)
        w = np.asarray(wilson, dtype=np.float64)
        e[: len(w)] = w
        out.append(e.tobytes())
    if h["pl_level"] > 0:
        e = np.zeros((h["polyakov_loop_size"], 2))
        p = np.asarray(poly, dtype=np.float64).reshape(-1, 2)
        e[: len(p)] = p
        out.append(e.tobytes())
    t = planes_from_links(np.asarray(links, dtype=np.float64), group, s["rowsize"], s["table_size"])
    out.append(t.astype(np.float32 if single else np.float64).tobytes())
    return b"".join(out)


def decode(data):
    """Sections of a file in the reference layout, plus what follows them."""
    w = np.frombuffer(data[: 4 * HEAD_WORDS], dtype=np.uint32)
    h = parse_header(w)
    group, dims = h["group"], h["full"]
    s = sizes(group, dims, h["iter"])
    pos = 4 * HEAD_WORDS
    out = {"header": h, "sizes": s}

    def take(n, dtype, shape):
        nonlocal pos
        a = np.frombuffer(data[pos:pos + n * np.dtype(dtype).itemsize], dtype=dtype).reshape(shape)
        pos += a.nbytes
        return a

    en = h["energies_size"]
    out["partials"] = take(2 * h["measurement_size_F"], np.float64, (-1, 2))
    out["actions"] = take(2 * en, np.float64, (-1, 2))
    if h["get_plaquettes"] or h["get_fmunu"] or _has_f(h):
        e = take(2 * h["energies_size_F"], np.float64, (-1, 2))
        out["plaquettes"] = e[:en]
        if _has_f(h):
            z = e[en:, 0] + 1j * e[en:, 1]
            out["fmunu"] = z.reshape(2, 3, en).transpose(2, 0, 1)   # [rec, generator, plane]
    if h["get_wilson"]:
        out["wilson"] = take(en, np.float64, (-1,))
    if h["pl_level"] > 0:
        out["polyakov"] = take(2 * h["polyakov_loop_size"], np.float64, (-1, 2))
    single = h["precision"] == 1
    table = take(4 * s["table_size"], np.float32 if single else np.float64, (-1, 4))
    out["table"] = table
    out["links"] = links_from_planes(table.astype(np.float64), group, s["V"], s["rowsize"])
    out["tail"] = data[pos:]
    out["table_end"] = pos
    return out
