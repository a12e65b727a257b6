"""The autotuner's table file (gr_tune_save / gr_tune_load / gr_tune_clear): host-only, no
device needed. Entries round-trip; a malformed or unknown entry rejects the whole file and
leaves the table as it was."""
import os

import pytest


def _lib():
    import granite_amd as ga

    try:
        return ga, ga.library()
    except Exception as exc:  # library not built here
        pytest.skip(f"library not available: {exc}")


def test_table_round_trip(tmp_path):
    ga, _ = _lib()
    ga.tune_clear()
    src = os.path.join(str(tmp_path), "in.txt")
    with open(src, "w") as f:
        f.write("# comment\n\n0 3 0 1\n1 2 0 11\n 2 1 0 2 \n0 0 1 0\n")
    ga.tune_load(src)
    out = os.path.join(str(tmp_path), "out.txt")
    ga.tune_save(out)
    rows = [ln.split() for ln in open(out) if ln.strip() and not ln.startswith("#")]
    assert sorted(map(tuple, rows)) == sorted([("0", "0", "1", "0"), ("0", "3", "0", "1"), ("1", "2", "0", "11"),
                                              ("2", "1", "0", "2")])
    ga.tune_clear()
    ga.tune_save(out)
    assert [ln for ln in open(out) if not ln.startswith("#")] == []


@pytest.mark.parametrize("bad", ["0 3 0 3\n", "1 2 0 3\n", "3 0 0 1\n", "0 4 0 1\n", "0 1 2 0\n", "0 1 0\n",
                                 "0 1 0 1 9\n", "x 1 0 1\n"])
def test_bad_entries_rejected(tmp_path, bad):
    ga, _ = _lib()
    ga.tune_clear()
    good = os.path.join(str(tmp_path), "good.txt")
    with open(good, "w") as f:
        f.write("1 1 0 2\n")
    ga.tune_load(good)
    path = os.path.join(str(tmp_path), "bad.txt")
    with open(path, "w") as f:
        f.write("2 2 0 4\n" + bad)
    with pytest.raises(ga.EngineError):
        ga.tune_load(path)
    out = os.path.join(str(tmp_path), "out.txt")
    ga.tune_save(out)
    rows = [ln.split() for ln in open(out) if not ln.startswith("#")]
    assert rows == [["1", "1", "0", "2"]]  # nothing of the rejected file was taken
    ga.tune_clear()


def test_missing_file_rejected(tmp_path):
    ga, _ = _lib()
    with pytest.raises(ga.EngineError):
        ga.tune_load(os.path.join(str(tmp_path), "absent.txt"))
