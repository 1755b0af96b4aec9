"""Variants files (mutant / perturbation sweeps; format in include/mbg.h), on the CPU: what each
statement does to the model's overrides, composition with a base model's own overrides, and
the error messages (with line numbers) for malformed files.  The reference has no
counterpart (README.md:33 lists mutant analysis as future work)."""
import subprocess

import numpy as np
import pytest

from tests.conftest import model_text
from upmabossg_amd import CLI_PATH, MbgError, Model

TEXT = """
// three variants of metastasis
variant wt { }
variant ko_p53 { p53.knockout; }
variant mix-1 {
    AKT1.rate_up *= 2.5;  AKT1.rate_up *= 2;     // factors compose: 5
    AKT2.rate_down *= 0.5e-1;
    CDH1.istate = 1;
    NICD.overexpress;
    sample_count = 500; seed_pseudorandom = 9;
}
"""


@pytest.fixture(scope="module")
def base():
    return Model(*model_text("metastasis"))


def test_statements_set_the_overrides(base):
    names = base.names()
    ix = {n: i for i, n in enumerate(names)}
    vs = base.variants(TEXT)
    assert [v[0] for v in vs] == ["wt", "ko_p53", "mix-1"]
    (_, wt, n0, s0), (_, ko, n1, s1), (_, mix, n2, s2) = vs
    assert (n0, s0, n1, s1, n2, s2) == (0, None, 0, None, 500, 9)
    up, down, ist = wt.get_variant()
    assert np.all(up == 1) and np.all(down == 1) and np.array_equal(ist, base.istate())
    up, down, ist = ko.get_variant()
    p = ix["p53"]
    assert up[p] == 0 and ist[p] == 0 and down[p] == 1
    assert np.count_nonzero(up != 1) == 1 and np.count_nonzero(ist != base.istate()) == 1
    up, down, ist = mix.get_variant()
    assert up[ix["AKT1"]] == np.float32(5.0)
    assert down[ix["AKT2"]] == np.float32(0.05)
    assert ist[ix["CDH1"]] == 1
    assert down[ix["NICD"]] == 0 and ist[ix["NICD"]] == 1
    assert np.count_nonzero(up != 1) == 1 and np.count_nonzero(down != 1) == 2
    # the base model itself is untouched
    up, down, ist = base.get_variant()
    assert np.all(up == 1) and np.all(down == 1)


def test_overrides_compose_with_the_base_models_own(base):
    b = base.clone()
    N = b.num_nodes
    up = np.full(N, 3.0, np.float32)
    ist = np.full(N, 0.25, np.float32)
    b.set_variant(mult_up=up, istate=ist)
    (_, v, _, _), = b.variants("variant x { p53.rate_up *= 2; AKT1.istate = 0.75; }")
    vu, vd, vi = v.get_variant()
    ix = {n: i for i, n in enumerate(b.names())}
    assert vu[ix["p53"]] == 6 and np.count_nonzero(vu == 3) == N - 1
    assert np.all(vd == 1)
    assert vi[ix["AKT1"]] == np.float32(0.75) and np.count_nonzero(vi == np.float32(0.25)) == N - 1


@pytest.mark.parametrize("text,msg", [
    ("variant a { foo.knockout; }", "line 1: unknown node 'foo'"),
    ("\nvariant a {\n p53.istate = 2; }", "line 3: istate must lie in [0, 1]"),
    ("variant a { p53.rate_up *= -1; }", "unexpected character '-'"),
    ("variant a { p53.rate_up *= 1e60; }", "rate factor must be finite and non-negative"),
    ("variant a {}\nvariant a {}", "line 2: duplicate variant name 'a'"),
    ("variant a { p53.knockout }", "expected ';'"),
    ("variant a { p53.knockout;", "unterminated block of variant 'a'"),
    ("varient a {}", "expected 'variant NAME {'"),
    ("variant a { sample_count = 0; }", "sample_count must be positive"),
    ("variant a { sample_count = 1.5; }", "expected a non-negative integer"),
    ("variant a { p53.foo; }", "unknown attribute 'foo'"),
    ("variant a { max_time = 3; }", "unknown setting 'max_time'"),
])
def test_errors_name_the_line(base, text, msg):
    with pytest.raises(MbgError) as e:
        base.variants(text)
    assert msg in str(e.value) and e.value.status == 1


def test_empty_file_has_no_variants(base):
    assert base.variants("// nothing\n") == []


def test_cli_reports_variant_file_errors(tmp_path):
    b, c = model_text("metastasis")
    (tmp_path / "m.bnd").write_text(b)
    (tmp_path / "m.cfg").write_text(c)
    (tmp_path / "v.txt").write_text("variant a { nosuch.knockout; }\n")
    r = subprocess.run([CLI_PATH, "-v", str(tmp_path / "v.txt"), str(tmp_path / "m.bnd"), str(tmp_path / "m.cfg")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "line 1: unknown node 'nosuch'" in r.stderr
    r = subprocess.run([CLI_PATH, "-v", str(tmp_path / "v.txt"), "-x", "p", "a", "b"], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage: MaBoSSG [-o prefix] bnd_file cfg_file" in r.stdout
