"""The MaBoSSG CLI end to end on the GPU: its <prefix>_probtraj.csv, <prefix>_fp.csv and stdout
text are compared byte for byte with text formatted here, independently, from the oracle's
tables of the same run (the GPU's tables equal the oracle's bit for bit, so any difference is
a formatting difference).

The formatter below restates the reference's writers:
  probtraj  src/statistics/window_average_small.cpp:181-250 (header :192-215, rows :219-248)
  stdout    src/statistics/final_states.cpp:53-66, src/statistics/fixed_states.cu:141-152,
            src/statistics/window_average_small.cpp:142-179
  fp        src/statistics/fixed_states.cu:154-187
  state     src/state.cpp:29-48 (" -- " between ON node names, "<nil>" when none)
with values printed as FP32 through the default ostream format (%g, 6 significant digits).
Normalisation is this build's: cells are exact fixed-point sums, P = cell 2^-fp / (n tick)
(counts / n in discrete time), TH = th 2^-ft / n (/ tick), evaluated in FP64 and printed as FP32.
The columns the reference leaves at 0 are filled by default (MBG_ERROR_COLUMNS=0 keeps the
zeros): ErrorProba / ErrorTH = sqrt((sum of squares 2^-e / n - mean^2) / (n - 1)), the standard
error of the mean over trajectories of each one's window fraction (window-average TH), from
the oracle's exact sums of squares; HD=0 = the window's total probability (the language has no
reference state, so every state is at Hamming distance 0).
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleModel
from tests.conftest import model_text
from upmabossg_amd import CLI_PATH

pytestmark = pytest.mark.gpu

_LIBM = ctypes.CDLL("libm.so.6")
_LIBM.log2f.restype = ctypes.c_float
_LIBM.log2f.argtypes = [ctypes.c_float]
f32 = np.float32


def g(v):
    return "%g" % float(f32(v))


def _se(n, mean, raw_sq, e):
    if n < 2:
        return 0.0
    var = math.ldexp(float(raw_sq), -e) / n - mean * mean
    return math.sqrt(var / (n - 1)) if var > 0 else 0.0


class Expected:
    def __init__(self, bnd, cfg, variant=None, n=None, seed=None, errors=True):
        self.errors = errors
        o = OracleModel(bnd, cfg)
        if variant is not None:
            o.set_variant(*variant)
        self.names = o.names()
        internal = o.internal()
        self.readouts = [i for i in range(len(self.names)) if not internal[i]]
        self.discrete = o.constant("discrete_time") != 0
        self.tick = f32(o.constant("time_tick"))
        self.n = n or int(o.constant("sample_count"))
        self.r = o.run(count=self.n, n_trajectories=self.n, seed=seed, threads=16, record_fixed=True)

    def proj(self, s):
        on = [self.names[self.readouts[j]] for j in range(len(self.readouts)) if (s >> j) & 1]
        return " -- ".join(on) if on else "<nil>"

    def full(self, words):
        on = [self.names[i] for i in range(len(self.names)) if (int(words[i >> 6]) >> (i & 63)) & 1]
        return " -- ".join(on) if on else "<nil>"

    def prob_d(self, raw):
        v = float(raw)
        if self.discrete:
            return v / self.n
        return math.ldexp(v, -self.r.prob_exponent) / (self.n * float(self.tick))

    def prob(self, raw):
        return f32(self.prob_d(raw))

    def th_d(self, w):
        v = math.ldexp(float(self.r.th[w]), -self.r.th_exponent) / self.n
        if not self.discrete:
            v /= float(self.tick)
        return v

    def th(self, w):
        return f32(self.th_d(w))

    def row(self, w):
        # the oracle's non-zero cells (window-major, states ascending): any number of readouts;
        # (state, P, ErrorProba) per state, the entropy and the total probability (HD=0)
        cw, cs, cv = self.r.cells
        lo, hi = np.searchsorted(cw, w, "left"), np.searchsorted(cw, w, "right")
        probs, ent, total = [], f32(0), 0.0
        for s, raw, sq in zip(cs[lo:hi], cv[lo:hi], self.r.cell_squares[lo:hi]):
            s = int(s)
            pd = self.prob_d(raw)
            p = f32(pd)
            if p == 0:
                continue
            ent = f32(ent + f32(f32(-_LIBM.log2f(p)) * p))
            total += pd
            probs.append((s, p, f32(_se(self.n, pd, sq, self.r.sq_exponent)) if self.errors else 0))
        return probs, ent, total

    def probtraj_csv(self):
        rows = [self.row(w) for w in range(self.r.windows)]
        out = "Time\tTH\tErrorTH\tH\tHD=0" + "\tState\tProba\tErrorProba" * max(len(p) for p, _, _ in rows) + "\n"
        for w, (probs, ent, total) in enumerate(rows):
            if self.errors:
                err_th = g(f32(_se(self.n, self.th_d(w), self.r.th_sq[w], self.r.th_sq_exponent)))
                hd0 = g(f32(total))
            else:
                err_th = hd0 = "0"
            out += "\t".join([g(f32(w) * self.tick), g(self.th(w)), err_th, g(ent), hd0])
            for s, p, e in probs:
                out += f"\t{self.proj(s)}\t{g(p)}\t{g(e)}"
            out += "\n"
        return out

    def fp_csv(self):
        F = len(self.r.fixed_counts)
        out = f"Fixed Points ({F})\n" + "FP\tProba\tState" + "".join("\t" + n for n in self.names) + "\n"
        for i in range(F):
            st = self.r.fixed_states[i]
            out += f"#{i}\t{g(f32(self.r.fixed_counts[i]) / f32(self.n))}\t{self.full(st)}"
            out += "".join("\t%d" % ((int(st[k >> 6]) >> (k & 63)) & 1) for k in range(len(self.names))) + "\n"
        return out

    def text(self):
        n = f32(self.n)
        out = "final points:\n"
        for s, c in zip(*self.r.final_cells):
            out += f"{g(f32(c) / n)} {self.proj(int(s))}\n"
        out += "fixed points:\n"
        for i, c in enumerate(self.r.fixed_counts):
            out += f"{g(f32(c) / n)} {self.full(self.r.fixed_states[i])}\n"
        for w in range(self.r.windows):
            probs, ent, _ = self.row(w)
            out += f"window ({g(f32(w) * self.tick)}, {g(f32(w + 1) * self.tick)}]\n"
            out += f"entropy: {g(ent)}\ntransition entropy: {g(self.th(w))}\n"
            for s, p, _ in probs:
                out += f"{g(p)} {self.proj(s)}\n"
        return out


CASES = {
    "metastasis": lambda: model_text("metastasis"),  # continuous, 1,000 trajectories, fixed points
    "cellcycle": lambda: (model_text("cellcycle")[0], model_text("cellcycle")[1] + "\nsample_count = 20000;\n"),
    "readouts24": lambda: _readouts24(),
    "ring100": lambda: _ring100(),
    "frozen": lambda: ("node A { logic = A; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }\n"
                       "node B { logic = !A; rate_up = @logic ? 2 : 0; rate_down = @logic ? 0 : 2; }\n",
                       "A.istate = 0.5; B.istate = 0.5; max_time = 5; time_tick = 1; sample_count = 1000;"),
}


def _ring100():
    # a ring model of the reference's generator (21 inputs per node: logic programs)
    from upmabossg_amd import synth

    return synth.ring_network(100, steps_to_fixed_point=50, sample_count=3000)


def _readouts24():
    # more than 20 non-internal nodes: sparse window cells (beyond the reference's limit)
    from upmabossg_amd import synth

    net = synth.random_network(48, 2, 3, seed=11, rates="loguniform", istate="uniform", readouts=24, max_time=10.0,
                               time_tick=0.5, sample_count=3000)
    return net.bnd(), net.cfg()


@pytest.mark.parametrize("name", sorted(CASES) + ["metastasis_no_error_columns"])
def test_cli_output_bytes(tmp_path, name):
    errors = not name.endswith("_no_error_columns")
    bnd, cfg = CASES[name.replace("_no_error_columns", "")]()
    bp, cp = tmp_path / "m.bnd", tmp_path / "m.cfg"
    bp.write_text(bnd)
    cp.write_text(cfg)
    print(f"case: CLI {name}", flush=True)
    env = dict(os.environ, MBG_ERROR_COLUMNS="1" if errors else "0")
    run = subprocess.run([CLI_PATH, "-o", str(tmp_path / "out"), str(bp), str(cp)], capture_output=True, text=True,
                         timeout=120, env=env)
    assert run.returncode == 0, run.stderr
    text = subprocess.run([CLI_PATH, str(bp), str(cp)], capture_output=True, text=True, timeout=120, env=env)
    assert text.returncode == 0, text.stderr
    want = Expected(bnd, cfg, errors=errors)
    assert (tmp_path / "out_probtraj.csv").read_text() == want.probtraj_csv()
    assert (tmp_path / "out_fp.csv").read_text() == want.fp_csv()
    assert text.stdout == want.text()
    if name == "metastasis":
        assert len(want.r.fixed_counts) > 0


VARIANTS = """
variant wt { }
variant ko_p53 { p53.knockout; }
variant oe_NICD { NICD.overexpress; AKT1.rate_up *= 2.5; }
variant mix { AKT2.rate_down *= 0.25; CDH1.istate = 1; TGFbeta.rate_up *= 0; sample_count = 700; seed_pseudorandom = 9; }
"""


def test_cli_variants_file_bytes(tmp_path):
    """-v: one sweep for every variant of the file; each variant's files and stdout block are
    byte for byte the formatted oracle tables of its model (the overrides applied), and the
    empty variant's files are the plain run's."""
    from upmabossg_amd import Model

    bnd, cfg = model_text("metastasis")
    bp, cp, vp = tmp_path / "m.bnd", tmp_path / "m.cfg", tmp_path / "v.txt"
    bp.write_text(bnd)
    cp.write_text(cfg)
    vp.write_text(VARIANTS)
    print("case: CLI variants metastasis", flush=True)
    run = subprocess.run([CLI_PATH, "-o", str(tmp_path / "out"), "-v", str(vp), str(bp), str(cp)],
                         capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    text = subprocess.run([CLI_PATH, "-v", str(vp), str(bp), str(cp)], capture_output=True, text=True, timeout=120)
    assert text.returncode == 0, text.stderr
    plain = subprocess.run([CLI_PATH, "-o", str(tmp_path / "plain"), str(bp), str(cp)], capture_output=True,
                           text=True, timeout=120)
    assert plain.returncode == 0, plain.stderr
    blocks = []
    for name, model, n, seed in Model(bnd, cfg).variants(VARIANTS):
        want = Expected(bnd, cfg, variant=model.get_variant(), n=n, seed=seed)
        assert (tmp_path / f"out_{name}_probtraj.csv").read_text() == want.probtraj_csv(), name
        assert (tmp_path / f"out_{name}_fp.csv").read_text() == want.fp_csv(), name
        blocks.append(f"variant {name}\n" + want.text())
    assert text.stdout == "".join(blocks)
    assert (tmp_path / "out_wt_probtraj.csv").read_text() == (tmp_path / "plain_probtraj.csv").read_text()
    assert (tmp_path / "out_wt_fp.csv").read_text() == (tmp_path / "plain_fp.csv").read_text()


def test_cli_variants_more_than_20_readouts(tmp_path):
    """-v on a model with 24 readouts (sparse window cells): the variants run one by one (sweeps
    keep dense tables), with the same counts, seeds and draws, and give the oracle's bytes."""
    from upmabossg_amd import Model

    bnd, cfg = _readouts24()
    text = "variant wt { }\nvariant ko { n3.knockout; n30.rate_up *= 2; sample_count = 1000; }\n"
    bp, cp, vp = tmp_path / "m.bnd", tmp_path / "m.cfg", tmp_path / "v.txt"
    bp.write_text(bnd)
    cp.write_text(cfg)
    vp.write_text(text)
    print("case: CLI variants readouts24", flush=True)
    run = subprocess.run([CLI_PATH, "-o", str(tmp_path / "out"), "-v", str(vp), str(bp), str(cp)],
                         capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    for name, model, n, seed in Model(bnd, cfg).variants(text):
        want = Expected(bnd, cfg, variant=model.get_variant(), n=n, seed=seed)
        assert (tmp_path / f"out_{name}_probtraj.csv").read_text() == want.probtraj_csv(), name
        assert (tmp_path / f"out_{name}_fp.csv").read_text() == want.fp_csv(), name
