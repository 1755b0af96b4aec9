"""libmbg.so on the CPU side (no GPU calls): exported symbols, front end, errors, codegen, CLI."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleModel
from tests.conftest import REFERENCE_MODELS, ROOT, model_text
from upmabossg_amd import CLI_PATH, LIB_PATH, MbgError, Model, synth


def _declared_symbols(header="mbg.h"):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(mbg_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mbg_[a-z_]+)$", out, re.M))
    assert exported == set(syms)


def test_mlp_library_exports_every_declared_symbol():
    from upmabossg_amd.mlp import MLP_LIB_PATH

    syms = _declared_symbols("mbg_mlp.h")
    assert "mbg_mlp_ensemble_predict" in syms and len(syms) >= 12
    lib = ctypes.CDLL(MLP_LIB_PATH)
    for s in syms:
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", MLP_LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mbg_[a-z_]+)$", out, re.M))
    assert exported == set(syms)


@pytest.mark.parametrize("name", REFERENCE_MODELS)
def test_front_end_matches_oracle_on_reference_models(name):
    b, c = model_text(name)
    m = Model(b, c)
    o = OracleModel(b, c)
    assert m.names() == o.names()
    assert np.array_equal(m.istate(), o.istate())
    assert np.array_equal(m.internal(), o.internal())
    assert not m.generic().any()
    for k in ["discrete_time", "sample_count", "max_time", "time_tick", "seed_pseudorandom"]:
        assert m.constant(k) == o.constant(k)


def test_front_end_matches_oracle_on_synthetic():
    for net in [synth.config_small(), synth.config_large(), synth.config_fixed_heavy(n=2000)]:
        m = Model(net.bnd(), net.cfg())
        o = OracleModel(net.bnd(), net.cfg())
        assert m.names() == o.names() == net.names
        assert np.array_equal(m.istate(), net.istate)
        assert np.array_equal(~m.internal(), np.isin(np.arange(net.n), net.readouts))


GOOD_BND = "node A { logic = !A; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }\n"


@pytest.mark.parametrize("bnd,cfg,status,msg", [
    ("node A { logic = A; rate_up = 1; }", "", 1, "rate_down"),
    ("node A { logic = A; rate_down = 1; }", "", 1, "rate_up"),
    (GOOD_BND + GOOD_BND, "", 1, "already exists"),
    ("nod A { logic = A; rate_up = 1; rate_down = 1; }", "", 1, "node"),
    (GOOD_BND, "B.istate = 1;", 1, "Unknown node"),
    (GOOD_BND, "max_time = $undefined;", 1, "unknown variable"),
    (GOOD_BND, "max_time = 5", 1, "syntax error"),
    (GOOD_BND, "max_time = 5 #;", 1, "invalid character"),
    ("node A { logic = B; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }", "", 1, "unknown node"),
    (GOOD_BND, "A.istate = 99999999999;", 1, "out of range"),
    ("node A { logic = A; rate_up = -1; rate_down = 1; }", "", 1, "non-negative"),
])
def test_parse_errors(bnd, cfg, status, msg):
    with pytest.raises(MbgError) as e:
        Model(bnd, cfg)
    assert e.value.status == status
    assert msg.lower() in str(e.value).lower()


def test_more_than_63_readouts_is_a_limit_error():
    # up to 20 readouts: dense window tables (the reference's limit, src/main.cpp:142-146);
    # 21 to 63: sparse tables; above: a limit error
    net = synth.random_network(64, 1, 2, seed=1, readouts=64)
    with pytest.raises(MbgError) as e:
        Model(net.bnd(), net.cfg())
    assert e.value.status == 2 and "63 non-internal" in str(e.value)
    net = synth.random_network(63, 1, 2, seed=1, readouts=63)
    assert Model(net.bnd(), net.cfg()).num_readouts == 63


def test_wide_fast_nodes_get_logic_programs():
    # the ring models of the reference's generator (data/generate-synth.py) read 2 sqrt(N) + 1
    # other nodes: above 20 a fast node's logic runs as a program instead of a truth table
    for n, steps in [(100, 0), (1000, 200)]:
        b, c = synth.ring_network(n, steps_to_fixed_point=steps, sample_count=100)
        m = Model(b, c)
        assert m.num_nodes == n
        assert "#define MBG_PROG 1" in m.generate_source()
    b, c = synth.config_small().bnd(), synth.config_small().cfg()
    assert "#define MBG_PROG 0" in Model(b, c).generate_source()
    ins = " & ".join(f"B{k}" for k in range(21))
    nodes = "".join(f"node B{k} {{ logic = B{k}; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }}\n"
                    for k in range(21))
    # operators that test more than zero-ness are refused for such nodes
    with pytest.raises(MbgError) as e:
        Model(nodes + f"node A {{ logic = ({ins}) + B0; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }}", "")
    assert e.value.status == 1 and "may then only use" in str(e.value)
    # generic rates keep their 16-input table limit
    with pytest.raises(MbgError) as e:
        Model(nodes + f"node A {{ logic = {ins}; rate_up = ({ins}) ? 2 : 1; rate_down = 1; }}", "")
    assert "limit 16" in str(e.value)


@pytest.mark.parametrize("case", ["ring100", "ring1000", "wide1", "wide2", "wide3", "small", "generic"])
def test_compiled_rates_equal_the_oracle_on_random_states(case):
    """The kernel's tables (truth tables, logic programs with clauses, generic rate tables),
    evaluated on the host, give the oracle's FP32 rate for every node at random states: this
    pins the logic compiler on the CPU (the oracle evaluates the expression trees directly)."""
    from tests.wide_models import wide_model

    if case.startswith("ring"):
        b, c = synth.ring_network(int(case[4:]), steps_to_fixed_point=100, sample_count=100)
    elif case.startswith("wide"):
        b, c = wide_model(seed=int(case[4:]))
    elif case == "small":
        net = synth.config_small()
        b, c = net.bnd(), net.cfg()
    else:
        b, c = EDGE_GENERIC
    m, o = Model(b, c), OracleModel(b, c)
    rng = np.random.default_rng(7)
    N = m.num_nodes
    for density in [0.1, 0.5, 0.9]:
        for _ in range(20 if N <= 100 else 3):
            bits = rng.random(N) < density
            words = o._words(bits)
            got = m.transition_rates(words)
            want = np.array([o.rate_f32(i, bits) for i in range(N)], np.float32)
            assert np.array_equal(got, want)


EDGE_GENERIC = ("node A { logic = B; rate_up = B ? 2.5 : (C ? 0.5 : 0); rate_down = A & !C ? 1 : 3; }\n"
                "node B { logic = A ^ C; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }\n"
                "node C { logic = !A | B; rate_up = @logic ? 1.5 : 0; rate_down = @logic ? 0 : 0.25; }\n", "")


def test_grammar_details():
    # keywords are case-insensitive, TRUE/FALSE, comments, precedence of ?: and unary minus
    bnd = ("Node A { logic = B and NOT C Or false; rate_up = @logic ? 2*3-4/2 : 0; rate_down = @logic ? 0 : -(-1.5e0); }\n"
           "NODE B { logic = A; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; } // trailing\n"
           "node C { logic = 1 ? A : B ? C : A; rate_up = @logic ? .5 : 0; rate_down = @logic ? 0 : 2.; }\n")
    cfg = "[A].istate = 3[1], 1[0];\nB.istate = TRUE;\n$v = 2 + 3 * 4;\nmax_time = $v;\n"
    m = Model(bnd, cfg)
    o = OracleModel(bnd, cfg)
    assert m.names() == ["A", "B", "C"] == o.names()
    assert m.istate()[0] == np.float32(0.75) == o.istate()[0]
    assert m.istate()[1] == 1.0
    assert m.constant("max_time") == 14.0 == o.constant("max_time")
    assert not m.generic().any()


def test_generated_source_is_specialised():
    net = synth.config_large()
    src = Model(net.bnd(), net.cfg()).generate_source()
    assert "#define MBG_NW 16" in src
    assert "#define MBG_DISCRETE 0" in src
    assert "mbg_simulate" in src and "mbg_philox" in src
    # launch-shape knobs are compile-time constants of the generated source
    assert "#define MBG_DEP_CAP " in src and "#define MBG_STATE_HBM " in src
    assert "mbg_init_readouts" in src  # the readout pass of models above 128 words
    b, c = model_text("cellcycle")
    src = Model(b, c).generate_source(trace=True)
    assert "#define MBG_NW 1" in src and "#define MBG_DISCRETE 1" in src and "#define MBG_TRACE 1" in src


def test_cli_usage_and_parse_errors(tmp_path):
    r = subprocess.run([CLI_PATH], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage: MaBoSSG [-o prefix] bnd_file cfg_file" in r.stdout
    r = subprocess.run([CLI_PATH, "-x", "a", "b", "c"], capture_output=True, text=True)
    assert r.returncode == 1
    bad = tmp_path / "bad.bnd"
    bad.write_text("node A {")
    cfg = tmp_path / "a.cfg"
    cfg.write_text("")
    r = subprocess.run([CLI_PATH, str(bad), str(cfg)], capture_output=True, text=True)
    assert r.returncode == 1 and "syntax error" in r.stderr
    r = subprocess.run([CLI_PATH, str(tmp_path / "missing.bnd"), str(cfg)], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr


def _gpu_visible():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="a GPU is visible: the no-GPU error path cannot be exercised")
def test_no_cpu_fallback_without_a_gpu(tmp_path):
    """Without a GPU the simulator and the surrogate network refuse to run (status 4, HIP error)
    instead of computing anything on the CPU."""
    from upmabossg_amd import Context
    from upmabossg_amd.mlp import Mlp, MlpError

    with pytest.raises(MbgError) as e:
        Context(0)
    assert e.value.status == 4
    with pytest.raises(MlpError):
        Mlp(10, 64, 1, 64, max_batch=64)
    net = synth.config_small()
    bnd, cfg = net.write(str(tmp_path / "m"))
    out = subprocess.run([CLI_PATH, bnd, cfg], capture_output=True, text=True)
    assert out.returncode == 1 and "HIP" in (out.stderr + out.stdout)
