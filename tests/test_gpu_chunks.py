"""Chunk-directory table I/O on the GPU (include/cdet_chunks.h).

Checks: a directory written by the test-side writer from oracle tables loads
bit for bit and evaluates bit-identically to the oracle; the GPU transform of
oracle samples matches the oracle's transformed tables; the library's own
directory reads back (test-side reader) as the device tables, with the
diagonal lines the reference's reduction uses; the chunked ordinary layout
agrees with the resident one under the index map and with the oracle's
empty-mask table (the reference's 3e-12 bound, tests/test_chunked_pipeline.cpp
:96-117); samples written by the GPU producer transform back to the same
tables; the reference's error texts for a missing READY, a manifest mismatch
and a short file."""
import os

import numpy as np
import pytest

import cdet_amd
import chunkdir
import helpers
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# tests/test_chunked_pipeline.cpp:134-151 (3x2, beta 2.25, mu 0.35, alpha 0.45/0.2,
# first_row), square Chebyshev grid of 8; node chunks of 7, sitepair chunks of 17.
P32 = dict(N1=3, N2=2, beta=2.25, U=2.7, mu=0.35, alpha_up=0.45, alpha_down=0.2, nc=8)
TEXTS = ("2.25", "0.35", "0.45", "0.2")


def ctx_of(s):
    p = cdet_amd.make_params(s["N1"], s["N2"], s["nc"], s["beta"], s["U"], s["mu"], s["alpha_up"], s["alpha_down"])
    return cdet_amd.Context(p), p


def chunked_oracle_tables(s, mask="first_row"):
    """Four kinds as the reference's chunked workflow defines them: twisted with
    the mask, ordinary with the empty mask, all replica blocks."""
    out = []
    for k in range(4):
        a = s["alpha_up"] if k in (0, 2) else s["alpha_down"]
        m = mask if k < 2 else "empty"
        out.append(O.Calc(s["N1"], s["N2"], s["beta"], s["mu"] - a, True, m, 0).build_table(s["nc"]))
    return out


def test_load_reference_layout_and_evaluate(gpu_available, tmp_path):
    s = helpers.SETS["c1_neq"]
    tabs = chunked_oracle_tables(s, s["mask"])
    print("writing 3x3 chunk directory", flush=True)
    chunkdir.write_directory(tmp_path, tabs, 3, 3, 64, 64, node_chunk=1000, sp_chunk=37,
                             texts=("4", "1", "1.15", "0.85"), mask="first_column")
    op = helpers.params_of("c1_neq")
    ctx = cdet_amd.Context(cdet_amd.make_params(3, 3, 64, s["beta"], s["U"], s["mu"], s["alpha_up"],
                                                s["alpha_down"],
                                                derived=(op.scale, op.diag_up_shift, op.diag_down_shift)))
    ctx.load_chunks(tmp_path, cdet_amd.lattice(True, "first_column", 0))
    for k in range(4):
        got = ctx.download_table(k)
        for g, w in zip(got, (tabs[k].upper, tabs[k].lower, tabs[k].ucorner, tabs[k].lcorner)):
            assert np.array_equal(g, w), k
    orc = O.Oracle(op, tabs)
    for verts in helpers.seeded_configs(20261203, [5, 7, 9]):
        assert np.array_equal(ctx.evaluate_all_subset_integrands(verts), orc.all_subsets(verts))
    ctx.close()


def test_transform_oracle_samples(gpu_available, tmp_path):
    s = P32
    samples = []
    for k in range(4):
        a = s["alpha_up"] if k in (0, 2) else s["alpha_down"]
        m = "first_row" if k < 2 else "empty"
        samples.append(O.Calc(3, 2, s["beta"], s["mu"] - a, True, m, 0).build_samples(8, 8))
    chunkdir.write_directory(tmp_path, None, 3, 2, 8, 8, node_chunk=7, sp_chunk=17, texts=TEXTS,
                             samples=samples, ready=False, coeffs=False)
    ctx, _ = ctx_of(s)
    ctx.transform_chunk_samples(tmp_path, cdet_amd.lattice(True, "first_row", 0))
    want = chunked_oracle_tables(s)
    for k in range(4):
        got = ctx.download_table(k)
        for g, w in zip(got, (want[k].upper, want[k].lower, want[k].ucorner, want[k].lcorner)):
            np.testing.assert_allclose(g, w, rtol=3e-12, atol=3e-12)
    ctx.close()


def test_write_read_back_and_layouts(gpu_available, tmp_path):
    s = P32
    ctx, p = ctx_of(s)
    ctx.build_tables(True, "first_row", 0, ordinary_layout=cdet_amd.ORDINARY_CHUNKED)
    spec = cdet_amd.chunk_spec(p, 7, 17, True, "first_row", 0)
    ctx.write_chunks(tmp_path, spec)
    assert open(tmp_path / "READY").read() == "ready\n"
    assert open(tmp_path / "manifest.txt").read().splitlines() == chunkdir.manifest_lines(
        3, 2, 8, 8, 7, 17, *TEXTS, True, "first_row", 0)
    dev = [ctx.download_table(k) for k in range(4)]
    for k in range(4):
        files = chunkdir.read_table(tmp_path, k, 3, 2, 8, 8)
        for g, w in zip(files, dev[k]):
            assert np.array_equal(g, w), k
    # diagonal lines: rows of the lower-branch coefficients, summed over v
    low = dev[0][1].reshape(144, 64)
    for (_, k, br, b, e) in chunkdir.read_tsv(tmp_path / "transform_jobs.tsv"):
        if br != 1:
            continue
        lines = chunkdir.read_diag(chunkdir.diag_file(tmp_path, k, b, e), 8)
        want_idx = [f for f in range(b, e) if (lambda d: d[0] == d[1] and d[2] == d[4] and d[3] == d[5])(
            chunkdir.decode(f, 3, 2))]
        assert sorted(lines) == want_idx
        if k == 0:
            for f, line in lines.items():
                np.testing.assert_allclose(line, low[f].reshape(8, 8).sum(axis=1), rtol=1e-14, atol=1e-15)
    # reload into a fresh context: bit for bit
    ctx2, _ = ctx_of(s)
    ctx2.load_chunks(tmp_path, cdet_amd.lattice(True, "first_row", 0))
    for k in range(4):
        for g, w in zip(ctx2.download_table(k), dev[k]):
            assert np.array_equal(g, w), k
    # chunked tables against the oracle's direct build (3e-12 bound of the reference's test)
    want = chunked_oracle_tables(s)
    for k in range(4):
        for g, w in zip(dev[k], (want[k].upper, want[k].lower, want[k].ucorner, want[k].lcorner)):
            np.testing.assert_allclose(g, w, rtol=3e-12, atol=3e-12)
    # resident ordinary layout = chunked (0,0) block at the site-major index
    ctx3, _ = ctx_of(s)
    ctx3.build_tables(True, "first_row", 0, ordinary_layout=cdet_amd.ORDINARY_RESIDENT)
    res = ctx3.download_table(2)[1].reshape(144, 64)
    chk = dev[2][1].reshape(144, 64)
    S = 6
    for x1 in range(3):
        for y1 in range(2):
            for x2 in range(3):
                for y2 in range(2):
                    e = (y1 * 3 + x1) * S + (y2 * 3 + x2)
                    f = (x1 * 2 + y1) * S + (x2 * 2 + y2)
                    assert np.array_equal(res[e], chk[f])
    with pytest.raises(RuntimeError, match="resident replica-\\(0,0\\) layout"):
        ctx3.write_chunks(tmp_path / "no", spec)
    for c in (ctx, ctx2, ctx3):
        c.close()


def test_gpu_samples_transform_to_built_tables(gpu_available, tmp_path):
    s = P32
    ctx, p = ctx_of(s)
    lat = cdet_amd.lattice(True, "first_row", 0, ordinary_layout=cdet_amd.ORDINARY_CHUNKED)
    cdet_amd.chunks_write_plan(tmp_path, cdet_amd.chunk_spec(p, 7, 17, True, "first_row", 0))
    ctx.write_chunk_samples(tmp_path, lat)
    # node-major file of the first job = column block of the oracle's samples
    orc = O.Calc(3, 2, s["beta"], s["mu"] - s["alpha_up"], True, "first_row", 0).build_samples(8, 8)
    first = np.fromfile(chunkdir.sample_file(tmp_path, 0, 0, 0, 7), dtype="<f8").reshape(7, 144)
    np.testing.assert_allclose(first, orc.upper.reshape(144, 64)[:, 0:7].T, rtol=1e-12, atol=1e-13)
    ctx2, _ = ctx_of(s)
    ctx2.transform_chunk_samples(tmp_path, lat)
    ctx.build_tables(True, "first_row", 0, ordinary_layout=cdet_amd.ORDINARY_CHUNKED)
    for k in range(4):
        for g, w in zip(ctx2.download_table(k), ctx.download_table(k)):
            assert np.array_equal(g, w), k
    ctx.close()
    ctx2.close()


def test_load_errors(gpu_available, tmp_path):
    s = P32
    ctx, p = ctx_of(s)
    ctx.build_tables(True, "first_row", 0, ordinary_layout=cdet_amd.ORDINARY_CHUNKED)
    ctx.write_chunks(tmp_path, cdet_amd.chunk_spec(p, 7, 17, True, "first_row", 0))
    with pytest.raises(RuntimeError, match="mismatch for SUBREGION_MASK"):
        ctx.load_chunks(tmp_path, cdet_amd.lattice(True, "checkerboard", 0))
    other, _ = ctx_of(dict(s, beta=2.5))
    with pytest.raises(RuntimeError, match="Chunk manifest mismatch for BETA: expected 2.5, found 2.25"):
        other.load_chunks(tmp_path)
    other.close()
    f = chunkdir.coeff_file(tmp_path, 1, 1, 17, 34)
    open(f, "ab").write(b"\0" * 8)
    with pytest.raises(RuntimeError, match="Unexpected file size for"):
        ctx.load_chunks(tmp_path)
    os.remove(tmp_path / "READY")
    with pytest.raises(RuntimeError, match="Chunked load-data directory is missing READY"):
        ctx.load_chunks(tmp_path)
    ctx.close()
