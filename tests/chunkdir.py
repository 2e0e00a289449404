"""Test-side reader and writer of the reference's chunk-directory layout
(include/cdet/chunked_load_data.h), written independently of the library so
that the library's files are checked against the layout the reference reads
and writes: paths :123-199, manifest :333-359, job plans :754-824, sample
files :947-996, corner files :998-1033, diagonal lines :1035-1093, coefficient
files :1095-1158."""
import os

import numpy as np

KINDS = ["twisted_up", "twisted_down", "ordinary_up", "ordinary_down"]
BRANCHES = ["upper", "lower"]


def token(b, e):
    return f"{b:010d}_{e:010d}"


def coeff_file(d, kind, br, b, e):
    return os.path.join(d, "coeffs", KINDS[kind], BRANCHES[br], f"sitepairs_{token(b, e)}.bin")


def sample_file(d, kind, br, b, e):
    return os.path.join(d, "samples", KINDS[kind], BRANCHES[br], f"nodes_{token(b, e)}.bin")


def corner_file(d, kind, br):
    return os.path.join(d, "corners", KINDS[kind], f"{BRANCHES[br]}.bin")


def diag_file(d, kind, b, e):
    return os.path.join(d, "diagonal_lines", KINDS[kind], f"sitepairs_{token(b, e)}.bin")


def manifest_lines(N1, N2, nu, nv, node_chunk, sp_chunk, beta, mu, a_up, a_dn, pbc, mask, radius):
    S = N1 * N2
    kv = {"layout_version": "1", "N1": str(N1), "N2": str(N2), "n_sites": str(S),
          "replica_sitepair_count": str(4 * S * S), "CHEB_NU": str(nu), "CHEB_NV": str(nv),
          "node_count": str(nu * nv), "node_chunk_size": str(node_chunk), "sitepair_chunk_size": str(sp_chunk),
          "BETA": beta, "MU": mu, "ALPHA_UP": a_up, "ALPHA_DOWN": a_dn, "PBC": "true" if pbc else "false",
          "SUBREGION_MASK": mask, "SUBREGION_SQUARE_RADIUS": str(radius),
          "LOAD_DATA_REAL": "double", "VECTOR_REAL": "double"}
    return [f"{k}={kv[k]}" for k in sorted(kv)]   # std::map order (byte order)


def plan(total, chunk):
    """(index, kind, branch, begin, end) for every kind, branch, range."""
    jobs, index = [], 1
    for k in range(4):
        for br in range(2):
            for b in range(0, total, chunk):
                jobs.append((index, k, br, b, min(total, b + chunk)))
                index += 1
    return jobs


def write_tsv(path, header, rows):
    with open(path, "w") as f:
        f.write(header)
        for r in rows:
            f.write("\t".join([str(r[0]), KINDS[r[1]], BRANCHES[r[2]]] + [str(x) for x in r[3:]]) + "\n")


def read_tsv(path):
    rows = []
    for line in open(path):
        line = line.rstrip("\n")
        if not line or line.startswith("#"):
            continue
        t = line.split("\t")
        rows.append(tuple([int(t[0]), KINDS.index(t[1]), BRANCHES.index(t[2])] + [int(x) for x in t[3:]]))
    return rows


def write_directory(d, tables, N1, N2, nu, nv, node_chunk, sp_chunk, texts, pbc=True, mask="first_row",
                    radius=0, samples=None, ready=True, coeffs=True):
    """Write a chunk directory: tables (four objects with upper/lower/ucorner/
    lcorner) as coeffs/ and corners/, optionally node-major samples/ from
    `samples` (four objects, [sitepair][node] arrays), and READY."""
    S = N1 * N2
    SP, NC = 4 * S * S, nu * nv
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "manifest.txt"), "w") as f:
        f.write("\n".join(manifest_lines(N1, N2, nu, nv, node_chunk, sp_chunk, *texts, pbc, mask, radius)) + "\n")
    nodes, trans = plan(NC, node_chunk), plan(SP, sp_chunk)
    write_tsv(os.path.join(d, "node_jobs.tsv"), "#index\ttable\tbranch\tnode_begin\tnode_end\n", nodes)
    write_tsv(os.path.join(d, "transform_jobs.tsv"), "#index\ttable\tbranch\tsitepair_begin\tsitepair_end\n", trans)
    write_tsv(os.path.join(d, "corner_jobs.tsv"), "#index\ttable\tbranch\n",
              [(1 + 2 * k + br, k, br) for k in range(4) for br in range(2)])
    for k in range(4):
        for br in range(2):
            t = tables[k] if tables is not None else samples[k]
            corner = t.ucorner if br == 0 else t.lcorner
            os.makedirs(os.path.dirname(corner_file(d, k, br)), exist_ok=True)
            np.asarray(corner, dtype="<f8").tofile(corner_file(d, k, br))
    if coeffs:
        for (_, k, br, b, e) in trans:
            t = tables[k]
            grid = (t.upper if br == 0 else t.lower).reshape(SP, NC)
            os.makedirs(os.path.dirname(coeff_file(d, k, br, b, e)), exist_ok=True)
            np.ascontiguousarray(grid[b:e], dtype="<f8").tofile(coeff_file(d, k, br, b, e))
    if samples is not None:
        for (_, k, br, b, e) in nodes:
            grid = (samples[k].upper if br == 0 else samples[k].lower).reshape(SP, NC)
            os.makedirs(os.path.dirname(sample_file(d, k, br, b, e)), exist_ok=True)
            np.ascontiguousarray(grid[:, b:e].T, dtype="<f8").tofile(sample_file(d, k, br, b, e))
    if ready:
        with open(os.path.join(d, "READY"), "w") as f:
            f.write("ready\n")


def read_table(d, kind, N1, N2, nu, nv):
    """(upper, lower, ucorner, lcorner) of one kind from coeffs/ and corners/."""
    S = N1 * N2
    SP, NC = 4 * S * S, nu * nv
    up, lo = np.zeros((SP, NC)), np.zeros((SP, NC))
    for (_, k, br, b, e) in read_tsv(os.path.join(d, "transform_jobs.tsv")):
        if k != kind:
            continue
        a = np.fromfile(coeff_file(d, k, br, b, e), dtype="<f8")
        assert a.size == (e - b) * NC
        (up if br == 0 else lo)[b:e] = a.reshape(e - b, NC)
    uc = np.fromfile(corner_file(d, kind, 0), dtype="<f8")
    lc = np.fromfile(corner_file(d, kind, 1), dtype="<f8")
    return up.ravel(), lo.ravel(), uc, lc


def read_diag(path, nu):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:8], dtype="<u8")[0])
    rec = 8 + 8 * nu
    assert len(raw) == 8 + n * rec
    out = {}
    for r in range(n):
        chunk = raw[8 + r * rec: 8 + (r + 1) * rec]
        out[int(np.frombuffer(chunk[:8], dtype="<u8")[0])] = np.frombuffer(chunk[8:], dtype="<f8").copy()
    return out


def decode(f, N1, N2):
    y2 = f % N2; f //= N2
    x2 = f % N1; f //= N1
    y1 = f % N2; f //= N2
    x1 = f % N1; f //= N1
    rb = f % 2; ra = f // 2
    return ra, rb, x1, y1, x2, y2
