"""The autotuner's table writer (gilaelectromagnetics.jl_amd/tune.py `table`):
a candidate replaces a table entry only when it beats the incumbent by more
than the margin, entries equal to the default are not written, and a results
file without records of some part leaves that part's tables as they are.
Runs on a scratch copy of csrc/gla_tune.hpp."""
import importlib.util
import json
import os
import re
import shutil

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gilaelectromagnetics.jl_amd")


@pytest.fixture()
def tune(tmp_path):
    spec = importlib.util.spec_from_file_location("gla_tune_tool", os.path.join(PKG, "tune.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    shutil.copy(os.path.join(PKG, "csrc", "gla_tune.hpp"), tmp_path / "gla_tune.hpp")
    mod.CSRC = str(tmp_path)
    return mod


def _block(src, tag):
    m = re.search(rf"// BEGIN TUNED {tag}\n(.*?)// END TUNED {tag}", src, flags=re.S)
    assert m, tag
    return m.group(1)


def _run(tune, tmp_path, recs, margin=0.05):
    path = tmp_path / "res.jsonl"
    path.write_text("\n".join(json.dumps(r) for r in recs))
    tune.table(str(path), margin)
    return (tmp_path / "gla_tune.hpp").read_text()


def test_margin_and_default(tune, tmp_path):
    before = (tmp_path / "gla_tune.hpp").read_text()
    recs = []
    for dt in ("c64", "c128"):
        # 96: parities together 20% faster -> entry 2; 128: 2% faster -> kept at
        # the default 1 (no entry); 160: slower -> default
        for n, t1, t2 in ((96, 30.0, 24.0), (128, 50.0, 49.0), (160, 80.0, 90.0)):
            for c, t in ((1, t1), (2, t2)):
                recs.append({"part": "xk", "dtype": dt, "n": n, "choice": c, "rows_us": [t / 2, 1, 1, 1, t / 2]})
    after = _run(tune, tmp_path, recs)
    for tag in ("XK C64", "XK C128"):
        body = _block(after, tag)
        assert "case 96: return 2;" in body
        assert "case 128:" not in body and "case 160:" not in body
    # every other table is byte-for-byte unchanged
    for tag in ("C64", "C128", "C64 SMALL", "C128 SMALL", "ROWS C64", "ROWS C128", "PLAN C64", "PLAN C128",
                "GG C64", "GG C128", "YK C64", "YK C128", "RELOAD C64", "RELOAD C128"):
        assert _block(after, tag) == _block(before, tag), tag


def test_incumbent_kept_within_margin(tune, tmp_path):
    # an existing entry stays unless the other choice is more than 5% faster
    recs = [{"part": "xk", "dtype": "c64", "n": 96, "choice": c, "rows_us": [t, 1, 1, 1, 0]}
            for c, t in ((1, 30.0), (2, 24.0))]
    first = _run(tune, tmp_path, recs)
    assert "case 96: return 2;" in _block(first, "XK C64")
    recs = [{"part": "xk", "dtype": "c64", "n": 96, "choice": c, "rows_us": [t, 1, 1, 1, 0]}
            for c, t in ((1, 29.0), (2, 30.0))]
    second = _run(tune, tmp_path, recs)
    assert "case 96: return 2;" in _block(second, "XK C64")


def test_single_value_parts(tune, tmp_path):
    # gg (default 7 / 3), yk (default 2), rl (default 0 / 2): only winners
    # beyond the margin that differ from the default are written
    recs = []
    for c, t in ((3, 100.0), (7, 80.0)):
        recs.append({"part": "gg", "dtype": "c128", "n": 320, "choice": c, "large_us": [0, 0, t, 0, 0]})
    for c, t in ((1, 90.0), (2, 100.0)):
        recs.append({"part": "yk", "dtype": "c128", "n": 512, "choice": c, "rows_us": [0, t / 2, 0, t / 2, 0]})
    for c, t in ((0, 50.0), (1, 49.0), (2, 70.0)):
        recs.append({"part": "rl", "dtype": "c64", "n": 64, "choice": c, "large_us": [0, 0, t, 0, 0]})
    after = _run(tune, tmp_path, recs)
    assert "case 320: return 7;" in _block(after, "GG C128")
    assert "case 512: return 1;" in _block(after, "YK C128")
    assert "case 64:" not in _block(after, "RELOAD C64")
