"""On-disk problem files (include/emagddm_problem.h): round trip and the reader's
validation (CPU only; the reader is host code)."""
import os

import numpy as np
import pytest

from emagddm_amd import workloads as W
from emagddm_amd.problem_io import read_problem, write_problem

FIELDS = ["xyz", "tet2node", "tet_material", "nu", "kappa", "tet_subdomain", "pec_pairs", "jq"]


@pytest.fixture(scope="module")
def prob():
    return W.make_problem(8, 2, materials="contrast", seed=5, name="micro_c")


def test_round_trip_is_exact(prob, tmp_path):
    path = str(tmp_path / "p.emfd")
    write_problem(path, prob)
    q = read_problem(path)
    for k in FIELDS:
        a, b = getattr(prob, k), getattr(q, k)
        assert a.shape == b.shape and np.array_equal(a, b), k
    assert q.n_sub == prob.n_sub and q.name == "p"
    # the layout of the header comment: magic, 40-byte header, padded arrays, checksum
    nn, nt, npc, nm = prob.n_nodes, prob.n_tets, len(prob.pec_pairs), prob.nu.size
    pad = lambda b: (b + 7) // 8 * 8
    size = 8 + 40 + pad(24 * nn) + pad(16 * nt) + 2 * pad(4 * nt) + 2 * pad(8 * nm) + \
        pad(8 * npc) + 96 * nt + 8
    assert os.path.getsize(path) == size


def test_without_source(prob, tmp_path):
    import dataclasses
    path = str(tmp_path / "nojq.emfd")
    write_problem(path, dataclasses.replace(prob, jq=None))
    assert read_problem(path).jq is None


def _corrupt(path, offset, value):
    with open(path, "r+b") as f:
        f.seek(offset)
        f.write(bytes([value]))


def test_reader_rejects_damage(prob, tmp_path):
    path = str(tmp_path / "p.emfd")
    write_problem(path, prob)
    raw = open(path, "rb").read()
    # a flipped payload byte -> checksum
    _corrupt(path, 200, raw[200] ^ 0x40)
    with pytest.raises(RuntimeError, match="checksum"):
        read_problem(path)
    # wrong magic
    open(path, "wb").write(b"NOTAFILE" + raw[8:])
    with pytest.raises(RuntimeError, match="not an EMAGFDP"):
        read_problem(path)
    # unsupported version
    open(path, "wb").write(raw[:8] + (7).to_bytes(4, "little") + raw[12:])
    with pytest.raises(RuntimeError, match="version 7"):
        read_problem(path)
    # truncated
    open(path, "wb").write(raw[:-100])
    with pytest.raises(RuntimeError, match="length"):
        read_problem(path)
    with pytest.raises(RuntimeError, match="cannot open"):
        read_problem(str(tmp_path / "missing.emfd"))


def test_reader_rejects_bad_indices(prob, tmp_path):
    import dataclasses
    path = str(tmp_path / "bad.emfd")
    t2n = prob.tet2node.copy()
    t2n[3, 1] = prob.n_nodes  # one past the last node
    write_problem(path, dataclasses.replace(prob, tet2node=t2n))
    with pytest.raises(RuntimeError, match="node index out of range"):
        read_problem(path)
    sub = prob.tet_subdomain.copy()
    sub[sub == 3] = 2  # subdomain 3 empty while n_sub = 8
    write_problem(path, dataclasses.replace(prob, tet_subdomain=sub))
    with pytest.raises(RuntimeError, match="subdomain 3 has no tets"):
        read_problem(path)
