"""The kernel assembly comparison (bayesian-rans-uq_amd/csrc/asm_diff.py):
bodies equal up to label numbering and comments compare identical, any
changed instruction is reported, and a pattern that matches no kernel is an
error."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "bayesian-rans-uq_amd", "csrc", "asm_diff.py")

BODY = """\t.text
_ZN2rq4fill{name}EPdld:                ; @_ZN2rq4fill{name}EPdld
; %bb.0:
\ts_load_dwordx2 s[0:1], s[4:5], 0x0   ; comment {tag}
\ts_cbranch_vccz .LBB{n}_2
.Lpost_getpc{m}:
\ts_add_u32 s98, s98, (.LBB{n}_2-.Lpost_getpc{m})&4294967295
.LBB{n}_2:
\t{last}
.Lfunc_end{n}:
_ZN2rq5otherEv:
\ts_endpgm
.Lfunc_end9:
"""


def _write(tmp_path, fname, **kw):
    p = tmp_path / fname
    p.write_text(BODY.format(**kw))
    return str(p)


def _run(a, b, pattern):
    return subprocess.run([sys.executable, SCRIPT, a, b, pattern], capture_output=True,
                          text=True)


def test_renumbered_labels_and_comments_compare_identical(tmp_path):
    a = _write(tmp_path, "a.s", name="Kern", n=3, m=10, tag="x", last="s_endpgm")
    b = _write(tmp_path, "b.s", name="Kern", n=7, m=42, tag="y", last="s_endpgm")
    r = _run(a, b, "fillKern")
    assert r.returncode == 0, r.stdout
    assert "identical" in r.stdout


def test_a_changed_instruction_is_reported(tmp_path):
    a = _write(tmp_path, "a.s", name="Kern", n=3, m=10, tag="x", last="s_endpgm")
    b = _write(tmp_path, "b.s", name="Kern", n=3, m=10, tag="x", last="s_nop 0")
    r = _run(a, b, "fillKern")
    assert r.returncode == 1
    assert "different" in r.stdout and "s_nop 0" in r.stdout


def test_a_pattern_without_a_kernel_is_an_error(tmp_path):
    a = _write(tmp_path, "a.s", name="Kern", n=3, m=10, tag="x", last="s_endpgm")
    r = _run(a, a, "no_such_kernel")
    assert r.returncode == 2
