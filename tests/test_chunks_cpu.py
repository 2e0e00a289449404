"""Chunk-directory plan files on the CPU: manifest text, job plans and their
exact cover (tests/test_chunked_pipeline.cpp:13-94, :134-157 settings:
3x2 lattice, CHEB 4x5, node chunks of 7, sitepair chunks of 17)."""
import os

import pytest

import cdet_amd
import chunkdir


def ref_spec(**over):
    s = cdet_amd.ChunkSpec(N1=3, N2=2, cheb_nu=4, cheb_nv=5, node_chunk_size=7, sitepair_chunk_size=17,
                           beta=2.25, mu=0.35, alpha_up=0.45, alpha_down=0.2, pbc=1, mask_kind=2,
                           subregion_square_radius=0)
    for k, v in over.items():
        setattr(s, k, v)
    return s


def test_manifest_text(tmp_path):
    cdet_amd.chunks_write_plan(tmp_path, ref_spec())
    got = open(tmp_path / "manifest.txt").read().splitlines()
    want = chunkdir.manifest_lines(3, 2, 4, 5, 7, 17, "2.25", "0.35", "0.45", "0.2", True, "first_row", 0)
    assert got == want


def cover(intervals, total):
    cur = 0
    for b, e in sorted(intervals):
        assert b == cur and b < e <= total
        cur = e
    assert cur == total


def test_job_plans_cover_exactly(tmp_path):
    cdet_amd.chunks_write_plan(tmp_path, ref_spec())
    nodes = chunkdir.read_tsv(tmp_path / "node_jobs.tsv")
    trans = chunkdir.read_tsv(tmp_path / "transform_jobs.tsv")
    corners = chunkdir.read_tsv(tmp_path / "corner_jobs.tsv")
    for jobs in (nodes, trans, corners):
        idx = [j[0] for j in jobs]
        assert len(set(idx)) == len(idx) and idx == list(range(1, len(idx) + 1))
    for k in range(4):
        for br in range(2):
            cover([(j[3], j[4]) for j in nodes if j[1] == k and j[2] == br], 20)
            cover([(j[3], j[4]) for j in trans if j[1] == k and j[2] == br], 4 * 36)
            assert sum(1 for j in corners if j[1] == k and j[2] == br) == 1
    # uneven last chunks, as the reference's test intends
    assert (nodes[2][3], nodes[2][4]) == (14, 20) and (trans[8][3], trans[8][4]) == (136, 144)
    assert open(tmp_path / "node_jobs.tsv").readline() == "#index\ttable\tbranch\tnode_begin\tnode_end\n"
    assert open(tmp_path / "corner_jobs.tsv").readline() == "#index\ttable\tbranch\n"
    assert nodes == chunkdir.plan(20, 7) and trans == chunkdir.plan(144, 17)


def test_spec_round_trip(tmp_path):
    cdet_amd.chunks_write_plan(tmp_path, ref_spec(mask_kind=4, subregion_square_radius=1, pbc=0))
    s = cdet_amd.chunks_read_spec(tmp_path)
    for f, _ in cdet_amd.ChunkSpec._fields_:
        assert getattr(s, f) == getattr(ref_spec(mask_kind=4, subregion_square_radius=1, pbc=0), f), f


def test_reference_manifest_aliases(tmp_path):
    # a manifest as the reference writes it, with a mask alias and upper-case reals
    lines = chunkdir.manifest_lines(3, 2, 4, 5, 7, 17, "2.25", "0.35", "0.45", "0.2", False, "Center-Square", 1)
    lines = [ln.replace("=double", "=Double") for ln in lines]
    os.makedirs(tmp_path, exist_ok=True)
    (tmp_path / "manifest.txt").write_text("# comment\n" + "\n".join(lines) + "\n")
    s = cdet_amd.chunks_read_spec(tmp_path)
    assert (s.mask_kind, s.subregion_square_radius, s.pbc, s.beta) == (4, 1, 0, 2.25)


@pytest.mark.parametrize("edit,msg", [
    (lambda t: t.replace("layout_version=1", "layout_version=2"), "mismatch for layout_version"),
    (lambda t: t.replace("VECTOR_REAL=double", "VECTOR_REAL=long_double"), "mismatch for VECTOR_REAL"),
    (lambda t: t.replace("node_count=20", "node_count=21"), "mismatch for node_count"),
    (lambda t: t.replace("BETA=2.25\n", ""), "missing BETA"),
    (lambda t: t.replace("SUBREGION_MASK=first_row", "SUBREGION_MASK=stripes"), "Unknown SUBREGION_MASK"),
])
def test_manifest_errors(tmp_path, edit, msg):
    cdet_amd.chunks_write_plan(tmp_path, ref_spec())
    p = tmp_path / "manifest.txt"
    p.write_text(edit(p.read_text()))
    with pytest.raises(RuntimeError, match=msg):
        cdet_amd.chunks_read_spec(tmp_path)


def test_bad_spec_rejected(tmp_path):
    with pytest.raises(RuntimeError, match="chunk sizes must be positive"):
        cdet_amd.chunks_write_plan(tmp_path, ref_spec(sitepair_chunk_size=0))
