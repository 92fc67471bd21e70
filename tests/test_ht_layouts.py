"""Hash-table layouts of the generated kernels against the reference table.

tests/native/ht_layouts.cpp compiles the device header ht_layouts.h as host
C++ and checks, for thousands of random key sets, filters and payload ranges,
that the compact and key-indexed layouts answer every probe exactly like the
reference's zero-filled `(k - kmin) % len` table (including its key-0 /
empty-slot match) and never write outside their words.  The second test
checks that every generated module builds each table in the layout its
probes read.
"""
import pathlib
import re
import subprocess

import pytest

from tests import queries as Q

ROOT = pathlib.Path(__file__).resolve().parents[1]
pa = pytest.importorskip("practice_amd")


def test_layouts_match_reference_table(tmp_path):
    exe = tmp_path / "ht_layouts"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "practice_amd" / "csrc"), "-o", str(exe),
                    str(ROOT / "tests" / "native" / "ht_layouts.cpp")], check=True)
    out = subprocess.run([str(exe), "3000"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok 3000 cases")


def _layout_calls(src):
    builds, probes = {}, {}
    for m in re.finditer(r"ht_insert_(\w+)\((?:\(int\*\))?P\.ht\[(\d+)\], k, ([^;]*)\);", src):
        builds[int(m.group(2))] = (m.group(1), m.group(3))
    for m in re.finditer(r"ht_probe_(compact(?:<\w+>)?|dense<\w+(?:, \w+)?>|ref)\((?:\(const int\*\))?(?:s_ht|P\.ht\[)(\d+)\]?, c\d+\[i\], ([^)]*)\)", src):
        probes[int(m.group(2))] = (m.group(1).split("<")[0], m.group(3))
    return builds, probes


def test_builds_and_probes_agree_on_layout():
    from tests.test_batch import _host_db, mixed_queries
    db = _host_db()
    for sql in mixed_queries():
        src = db.generate_code(sql)
        builds, probes = _layout_calls(src)
        assert set(builds) == set(probes), sql
        for j in builds:
            kind_b, args_b = builds[j]
            kind_p, args_p = probes[j]
            assert kind_b == kind_p, (sql, j, kind_b, kind_p)
            # same table geometry (key range / width / payload base) on both sides
            nb = [a.strip() for a in args_b.split(",")]
            np_ = [a.strip() for a in args_p.split(",")]
            if kind_b == "dense":
                assert nb[:4] == np_[:4], (sql, j, nb, np_)
            elif kind_b == "compact":
                assert nb[:6] == np_[:6], (sql, j, nb, np_)
