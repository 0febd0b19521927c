"""Host-side CLI logic (input sniffing, mask validation) -- reference cli.py:45-124."""
import json

import numpy as np
import pytest

from contrafold_gpu_amd import cli


def test_literal_sequence():
    recs, from_file = cli.load_records("ggggaaaacccc")
    assert recs == [(0, "GGGGAAAACCCC", None)] and not from_file
    with pytest.raises(SystemExit):
        cli.load_records("not-a-sequence!")


def test_jsonl_fasta_lines(tmp_path):
    p = tmp_path / "a.jsonl"
    p.write_text(json.dumps({"id": "r1", "seq": "acgu", "constrain": "0010"}) + "\n\n" +
                 json.dumps({"seq": "GGGAAACCC"}) + "\n")
    recs, ff = cli.load_records(str(p))
    assert ff and recs[0] == ("r1", "ACGU", "0010") and recs[1][1] == "GGGAAACCC" and recs[1][0] == 2
    f = tmp_path / "a.fa"
    f.write_text(">x\nACG\nUU\n>y\nGGG\n")
    recs, _ = cli.load_records(str(f))
    assert recs == [("x", "ACGUU", None), ("y", "GGG", None)]
    t = tmp_path / "a.txt"
    t.write_text("acgu\n\nggg\n")
    recs, _ = cli.load_records(str(t))
    assert [r[1] for r in recs] == ["ACGU", "GGG"]
    bad = tmp_path / "b.jsonl"
    bad.write_text('{"id": 1}\n')
    with pytest.raises(SystemExit):
        cli.load_records(str(bad))


def test_build_mask():
    assert cli.build_mask("ACGU", None) is None
    assert list(cli.build_mask("ACGU", "0101")) == [0, 1, 0, 1]
    assert list(cli.build_mask("ACG", [1, 0, 0])) == [1, 0, 0]
    with pytest.raises(SystemExit):
        cli.build_mask("ACGU", "01")
    with pytest.raises(SystemExit):
        cli.build_mask("ACG", [0, 2, 0])
