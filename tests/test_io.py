"""CPU checks of the text formats at the boundary (no GPU): the SMS loader
(src/matrices.rs:662-732) against the reference's own loader output for
cuda_test/matrix.txt, the .wdm writer (src/wdm_files.rs:132-210) byte for
byte against files the reference's writer produced, the .wdm reader
(src/wdm_files.rs:20-128) on those files and on malformed ones, and the
number of entries a resumed sequence loop appends (src/main.rs:105-131)."""
import os

import numpy as np
import pytest

from rust_rank_amd import engine

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = 2013265921


def test_sms_matches_reference_loader_on_matrix_txt():
    A = engine.load_sms(os.path.join(G, "matrix.sms"), P)  # cuda_test/matrix.txt
    g = np.load(os.path.join(G, "matrix_txt.npz"))
    assert (A.m, A.n) == tuple(int(x) for x in g["shape"])
    assert np.array_equal(A.row_ptr.astype(np.int64), g["row_ptr"].astype(np.int64))
    assert np.array_equal(A.col.astype(np.int64), g["col"].astype(np.int64))
    assert np.array_equal(A.val, g["val"])


def _sms(tmp_path, text, p=97):
    f = tmp_path / "a.sms"
    f.write_text(text)
    return engine.load_sms(str(f), p)


def test_sms_parsing_rules(tmp_path):
    A = _sms(tmp_path, "% comment\n\n3 4 M\n1 2 -1\n1 4 100\n 2 1 +5 \n3 3 0\n1 2\n0 0 0\n")
    assert (A.m, A.n) == (3, 4)
    assert A.row_ptr.tolist() == [0, 2, 3, 4]
    assert A.col.tolist() == [1, 3, 0, 2]
    assert A.val.tolist() == [96, 3, 5, 0]  # -1 -> p - 1, 100 mod 97, explicit zero kept
    # empty trailing rows, and an entry of an earlier row joining the open row (as the reference does)
    B = _sms(tmp_path, "4 3 x\n2 1 7\n1 3 8\n")
    assert B.row_ptr.tolist() == [0, 0, 2, 2, 2] and B.col.tolist() == [0, 2]


@pytest.mark.parametrize("text", ["2 2 M\n1 3 1\n", "2 2 M\n3 1 1\n", "2 2 M\n1 1 x\n", "2 2 M\n1 1 4294967296\n",
                                  "2 2 M\n-1 1 1\n"])
def test_sms_rejects_bad_entries(tmp_path, text):
    with pytest.raises(engine.RankEngineError):
        _sms(tmp_path, text)


def test_wdm_writer_matches_reference_bytes(tmp_path):
    g = np.load(os.path.join(G, "wdm_ref.npz"))
    i = 0
    while f"meta{i}" in g:
        p, m, n, b, ln = (int(x) for x in g[f"meta{i}"])
        path = str(tmp_path / f"w{i}.wdm")
        engine.wdm_save(path, m, n, p, g[f"dr{i}"], g[f"dc{i}"], g[f"V{i}"], g[f"curv{i}"], g[f"seq{i}"])
        assert open(path, "rb").read() == bytes(g[f"file{i}"]), i
        i += 1
    assert i == 3


def test_wdm_reader_on_reference_files(tmp_path):
    g = np.load(os.path.join(G, "wdm_ref.npz"))
    for i in range(3):
        p, m, n, b, ln = (int(x) for x in g[f"meta{i}"])
        path = tmp_path / f"r{i}.wdm"
        path.write_bytes(bytes(g[f"file{i}"]))
        w = engine.wdm_load(str(path))
        assert (w["m"], w["n"], w["p"], w["b"], w["len"]) == (m, n, p, b, ln)
        for k, gk in (("row_precond", "dr"), ("col_precond", "dc"), ("V", "V"), ("curv", "curv"), ("seq", "seq")):
            assert np.array_equal(w[k], g[f"{gk}{i}"]), (i, k)


def _lines(tmp_path, text):
    f = tmp_path / "x.wdm"
    f.write_text(text)
    return str(f)


def test_wdm_reader_errors_and_extra_tokens(tmp_path):
    good = "2 3 97 2 1\n1 2\n3 4 5\n6 7 8\n9 10 11\n12 13 99\n"
    w = engine.wdm_load(_lines(tmp_path, good))  # sequence line: its first Nlen values only
    assert w["seq"].ravel().tolist() == [12, 13]
    bad = {
        "Missing num_v": "2 3 97 2\n",
        "v vector length does not match matrix columns": "2 3 97 2 1\n1 2\n3 4 5\n6 7\n9 10 11\n12 13\n",
        "curv vector length does not match matrix columns": "2 3 97 2 1\n1 2\n3 4 5\n6 7 8\n9 10\n12 13\n",
        "seq vector length does not match expected sequence length": "2 3 97 2 1\n1 2\n3 4 5\n6 7 8\n9 10 11\n12\n",
        "Row preconditioner length does not match matrix rows": "2 3 97 2 1\n1\n3 4 5\n6 7 8\n9 10 11\n12 13\n",
    }
    for msg, text in bad.items():
        with pytest.raises(engine.RankEngineError) as e:
            engine.wdm_load(_lines(tmp_path, text))
        assert msg in str(e.value), (msg, str(e.value))


def test_rank_from_wdm_refusals_before_the_device(tmp_path):
    """rr_rank_from_wdm checks the file before any device work: a missing
    file, a b or prime it cannot use, an order the sequence does not reach,
    and sizes no host can hold are errors (status and message), not crashes."""
    def write(name, text):
        f = tmp_path / name
        f.write_text(text)
        return str(f)
    good = "2 3 97 5 1\n1 2\n3 4 5\n6 7 8\n9 10 11\n1 2 3 4 5\n"
    cases = [
        (str(tmp_path / "missing.wdm"), "cannot open"),
        (write("p91.wdm", good.replace(" 97 ", " 91 ")), "prime"),
        (write("huge.wdm", "4611686018427387904 3 97 5 1\n"), "host allocation"),
    ]
    for path, msg in cases:
        with pytest.raises(engine.RankEngineError, match=msg):
            engine.rank_from_wdm(path)
    path = write("good.wdm", good)
    for order in (5, 9):
        with pytest.raises(engine.RankEngineError, match="order"):
            engine.rank_from_wdm(path, order=order)


def _reference_loop_length(have, max_nlen):
    """src/main.rs:105-131 replayed: returns the sequence length after the loop."""
    length = have
    to_be_produced = max_nlen - length  # computed before S_0 is appended
    if length == 0:
        length += 1
    for _ in range(to_be_produced // 2):
        length += 2
    return length


def test_resume_count_matches_reference_loop():
    for max_nlen in (1, 2, 7, 8, 9, 64, 65, 1025):
        for have in range(0, max_nlen + 1):
            got = engine.seq_resume_count(have, max_nlen)
            assert have + got == _reference_loop_length(have, max_nlen), (have, max_nlen)
    # a fresh reference-policy sequence is rr_seq_len's length
    for m, n, b in ((2000, 1800, 8), (500, 400, 4), (100, 90, 1)):
        L = engine.seq_len(m, n, b, reference_policy=True)
        assert engine.seq_resume_count(0, L - 1) == L
